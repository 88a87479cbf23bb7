// Launch wrappers of the heterodyned-likelihood kernels (qcbc_het.hip).
// Host-safe: no device code in here.
#pragma once

#include <hip/hip_runtime_api.h>

#include "qcbc_types.h"

namespace qcbc {

constexpr int HET_MAX_J = 8;  // Legendre order per region (the reference uses 2)
constexpr int HET_MAX_J1 = HET_MAX_J + 1;
constexpr int HET_SETUP_BLOCKS = 64;  // blocks of the per-bin set-up pass (partial sums of logLb)

// fullphaseamp of point 0 on n frequencies (freq null: f_i = i/Tobs, f_0 = 1/Tobs);
// amp, phs: nd rows of n
hipError_t launch_phaseamp(const Point* pts, const double* freq, int n, int nd, double Tobs, double sqT, double* amp,
                           double* phs, hipStream_t s);
// heterodyne at point 0 from its h22 amplitude/phase wa, wp on the N/2 bins.
// bins: 3 nd N/2 doubles of scratch, partial: HET_SETUP_BLOCKS 2 nd doubles.
hipError_t launch_het_setup(const Point* pts, const double* wa, const double* wp, const double* D, const double* SN,
                            int nd, int N, double Tobs, double sqT, const int* fgrid, int M, int J, int MN, int MM,
                            double* bins, double* partial, double* logLb, double* freq, double* amp, double* phs,
                            double* IP, double* SL, double* lc, double* ls, double* rc, double* rs, double* aa,
                            double* w, double* wf, double* pl, hipStream_t s);
// the sample powers pl (4 x M: f^(1/6), f^(-1/6), 1/f, log f) of M frequencies
hipError_t launch_het_planes(const double* freq, int M, double* pl, hipStream_t s);
// evaluation weights w and wf (3 nd M each) and sample powers pl (4 M) from
// imported set-up arrays
hipError_t launch_het_weights(const double* freq, const double* amp, const double* IP, const double* SL,
                              const double* lc, const double* ls, const double* rc, const double* rs,
                              const double* aa, int M, int J, int nd, double* w, double* wf, double* pl,
                              hipStream_t s);
// waves of the evaluation kernel per point (1, 2 or 4) for B points on a grid of M samples
int het_waves_per_point(int M, int B);
// log_likelihood_het for B set-up points
hipError_t launch_het_ll(const Point* pts, int B, int nd, const double* freq, const double* pl, const double* ha,
                         const double* hp, const double* wf, int M, double logLb, double* logL,
                         hipStream_t s);

}  // namespace qcbc
