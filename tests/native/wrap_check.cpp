// Checks qcbc::wrap_down / wrap_up against the literal loops of update()
// (QuickCBC.c:2336-2342) bit for bit, for t = 2 pi and pi, on random angles
// up to 1e4 rad, on values at and next to every binade boundary and on
// multiples of t.  Prints "ok <count>" or the first mismatch and exits 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "qcbc_wrap.h"

static double lit_down(double y, double t) {
  while (y > t) y -= t;
  return y;
}
static double lit_up(double y, double t) {
  while (y < 0.0) y += t;
  return y;
}
static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

static long long n_checked = 0;
static bool check(double y, double t) {
  n_checked++;
  const double a = qcbc::wrap_down(y, t), b = lit_down(y, t);
  const double c = qcbc::wrap_up(y, t), d = lit_up(y, t);
  if (!same(a, b) || !same(c, d)) {
    std::printf("mismatch y=%.17g t=%.17g: down %.17g vs %.17g, up %.17g vs %.17g\n", y, t, a, b, c, d);
    return false;
  }
  return true;
}

int main() {
  const double ts[2] = {6.283185307179586476925286766559, 3.141592653589793238462643383279};
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(-1.0e4, 1.0e4), L(-3.0, 4.0), S(0.0, 1.0);
  for (double t : ts) {
    for (int i = 0; i < 200000; i++)
      if (!check(U(rng), t)) return 1;
    for (int i = 0; i < 200000; i++) {  // log-uniform magnitudes, both signs
      const double m = std::pow(10.0, L(rng));
      if (!check(S(rng) < 0.5 ? m : -m, t)) return 1;
    }
    for (int e = -4; e <= 14; e++) {  // binade boundaries and their neighbours
      const double p = std::ldexp(1.0, e);
      double v = p;
      for (int k = 0; k < 40; k++) v = std::nextafter(v, 0.0);
      for (int k = 0; k < 80; k++, v = std::nextafter(v, 1e300))
        if (!check(v, t) || !check(-v, t) || !check(v + t, t) || !check(-v - t, t)) return 1;
    }
    for (int k = -1600; k <= 1600; k++) {  // multiples of t and their neighbours
      const double v = k * t;
      if (!check(v, t) || !check(std::nextafter(v, 1e300), t) || !check(std::nextafter(v, -1e300), t)) return 1;
    }
    if (!check(t, t) || !check(0.0, t) || !check(-0.0, t) || !check(1.0e4, t) || !check(-1.0e4, t)) return 1;
  }
  std::printf("ok %lld\n", n_checked);
  return 0;
}
