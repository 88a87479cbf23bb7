// Angle wrapping with the result of update()'s loops (QuickCBC.c:2336-2342),
//   while (y > t) y -= t;   while (y < 0) y += t;
// step for step, in far fewer steps for a large |y|.  Header-only, for the
// device code and for a host build of the same functions (the CPU test runs
// them against the literal loops; compile with -ffp-contract=off).
//
// Why a run of steps can be taken at once: while y and y - t lie in one
// binade [p, 2p) (ulp u), every y there is a multiple of u and
// fl(y - t) = y - c with the same c for every such y: c = t rounded to a
// multiple of u when t is not half-way between two of them; in the one binade
// where it is (u = twice t's lowest set bit) every rounded result has an even
// last bit, so from the first step in that binade on each step subtracts the
// same one of the two neighbours.  Two literal steps y -> y1 -> y2 inside one
// binade therefore give c = y1 - y2 exactly, and k more steps give
// y2 - k c exactly (a multiple of u in [p, 2p)) as long as each step's exact
// difference stays at or above p; k is chosen one or two short of that limit
// and the literal loop takes the binade's last few steps.  Steps that add t to
// a negative y mirror these (round-to-nearest-even is symmetric).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define QCBC_WRAP_HD __host__ __device__
#else
#define QCBC_WRAP_HD
#endif

namespace qcbc {

// while (y > t) y -= t;
QCBC_WRAP_HD inline double wrap_down(double y, double t) {
  while (y > t) {
    y -= t;
    if (y > 16.0 * t) {
      const double y2 = y - t;
      int e1, e2;
      (void)frexp(y, &e1);
      (void)frexp(y2, &e2);
      if (e1 == e2) {
        const double c = y - y2;          // exact: both multiples of the binade's ulp
        const double p = ldexp(0.5, e2);  // the binade's lower end
        const double k = floor((y2 - p - t - c) / c) - 1.0;
        y = k > 0.0 ? y2 - k * c : y2;  // k c and the difference are exact
      } else {
        y = y2;
      }
    }
  }
  return y;
}

// while (y < 0) y += t;
QCBC_WRAP_HD inline double wrap_up(double y, double t) {
  while (y < 0.0) {
    y += t;
    if (y < -16.0 * t) {
      const double y2 = y + t;
      int e1, e2;
      (void)frexp(y, &e1);
      (void)frexp(y2, &e2);
      if (e1 == e2) {
        const double c = y2 - y;
        const double p = ldexp(0.5, e2);
        const double k = floor((-y2 - p - t - c) / c) - 1.0;
        y = k > 0.0 ? y2 + k * c : y2;
      } else {
        y = y2;
      }
    }
  }
  return y;
}

}  // namespace qcbc
