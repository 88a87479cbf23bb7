// Launch wrappers of the data-conditioning kernels (qcbc_condition.hip).
#pragma once

#include <hip/hip_runtime_api.h>

namespace qcbc {

// tukey() (Utilities.c:635-654) in place on nd rows of N samples
hipError_t launch_tukey(double* x, int nd, int N, double alpha, hipStream_t s);
// nd rows of N/2 + 1 complex coefficients (double2) into GSL's half-complex
// layout, nd rows of N, times fac
hipError_t launch_pack_halfcomplex(const void* X, int nd, int N, double fac, double* D, hipStream_t s);
// main's spectra (QuickCBC.c:286-306) of nd rows of N/2 + 1 unscaled
// coefficients X (double2): pspec, wcheck (nd x N/2, each may be null) and
// the whitened coefficients W (nd x (N/2 + 1) double2, whiten() of X)
hipError_t launch_spectra(const void* X, const double* SN, int nd, int N, double Tobs, void* W, double* pspec,
                          double* wcheck, hipStream_t s);
// the whitened series DW (nd x N) from the unnormalised backward transform w
// of W: w / N / (sqrt(N/2)/sqrt(Tobs)) (QuickCBC.c:270-276)
hipError_t launch_whitened_series(const double* w, int nd, int N, double Tobs, double* DW, hipStream_t s);
// whiten() (Utilities.c:294-309) in place on one half-complex row of N
hipError_t launch_whiten(double* data, const double* Sn, int N, hipStream_t s);
// tukey_scale() (Utilities.c:656-683): s12 = {mean, rms} of the window
hipError_t launch_tukey_scale(int N, double alpha, double* s12, hipStream_t s);
// makespec (Utilities.c:1908-1959): the n knots (xa, ya = ln S), lines = 5
// rows of nlines (frequency, height, Q, width, deltafmax); c: n, w: 3 n
// scratch; SN: N/2
hipError_t launch_makespec(const double* xa, const double* ya, int n, const double* lines, int nlines, double Tobs,
                           int N, double* c, double* w, double* SN, hipStream_t s);

}  // namespace qcbc
