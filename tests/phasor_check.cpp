// Host check of the device phasor (ddf_device.h: phasor): |exp(i phase) - phasor(phase)| against
// long-double sin / cos over phases of 1e-3 .. 1e8 rad and around the quadrant boundaries.
// Built and run by test_phasor.py; exits 1 above 3e-7.
#include "ddf_device.h"
#include <cstdio>
#include <cmath>
#include <random>
using namespace ddf;
int main() {
  std::mt19937_64 g(1);
  double worst = 0, wphase = 0; double sum2 = 0; long n = 0;
  for (int e = -3; e <= 8; ++e) {
    std::uniform_real_distribution<double> U(-std::pow(10.0, e), std::pow(10.0, e));
    for (int i = 0; i < 200000; ++i) {
      double ph = U(g);
      cd z = phasor(ph);
      long double c = cosl((long double)ph), s = sinl((long double)ph);
      double err = std::sqrt(double((z.re - c) * (z.re - c) + (z.im - s) * (z.im - s)));
      if (err > worst) { worst = err; wphase = ph; }
      sum2 += err * err; ++n;
    }
  }
  // quadrant boundaries
  for (long q : {-8L, -7L, -6L, -5L, -4L, -3L, -2L, -1L, 0L, 1L, 2L, 3L, 4L, 5L, 6L, 7L, 8L, 1000001L, -1000001L, 40000003L, -40000002L})
   for (double d : {-1e-9, 0.0, 1e-9, 0.785398, -0.785398}) {
    double ph = q * 1.5707963267948966 + d;
    cd z = phasor(ph);
    long double c = cosl((long double)ph), s = sinl((long double)ph);
    double err = std::sqrt(double((z.re - c) * (z.re - c) + (z.im - s) * (z.im - s)));
    if (err > worst) { worst = err; wphase = ph; }
  }
  std::printf("max abs err %.3e at phase %.6g, rms %.3e\n", worst, wphase, std::sqrt(sum2 / n));
  return worst < 3e-7 ? 0 : 1;
}
