// Launch wrappers of the data-conditioning kernels (qcbc_condition.hip).
#pragma once

#include <hip/hip_runtime_api.h>

namespace qcbc {

// tukey() (Utilities.c:635-654) in place on nd rows of N samples
hipError_t launch_tukey(double* x, int nd, int N, double alpha, hipStream_t s);
// nd rows of N/2 + 1 complex coefficients (double2) into GSL's half-complex
// layout, nd rows of N, times fac
hipError_t launch_pack_halfcomplex(const void* X, int nd, int N, double fac, double* D, hipStream_t s);
// makespec (Utilities.c:1908-1959): the n knots (xa, ya = ln S), lines = 5
// rows of nlines (frequency, height, Q, width, deltafmax); c: n, w: 3 n
// scratch; SN: N/2
hipError_t launch_makespec(const double* xa, const double* ya, int n, const double* lines, int nlines, double Tobs,
                           int N, double* c, double* w, double* SN, hipStream_t s);

}  // namespace qcbc
