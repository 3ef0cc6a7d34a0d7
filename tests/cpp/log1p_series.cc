// Host check of akx_log1p<float> (akx_math.hpp): the Taylor series used for
// |x| < 1/16 and the library call beyond, against log1p in double, over a
// sweep of x through both branches and across the branch point.  Prints the
// largest relative error (in units of FP32 epsilon) of each branch and the
// number of points.
#include "../../akantu2_amd/csrc/akx_math.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>

int main() {
  double worst_series = 0., worst_lib = 0.;
  long n = 0;
  for (long k = -400000; k <= 400000; ++k) {
    const float x = float(k) * 2.5e-7f; // [-0.1, 0.1]
    if (x == 0.f)
      continue;
    const double ref = std::log1p(double(x));
    const double got = double(akx::akx_log1p<float>(x));
    const double rel = std::fabs(got - ref) / std::fabs(ref) / FLT_EPSILON;
    if (std::fabs(x) < 0.0625f)
      worst_series = std::fmax(worst_series, rel);
    else
      worst_lib = std::fmax(worst_lib, rel);
    ++n;
  }
  // tiny arguments: log1p(x) = x to FP32 precision
  for (int e = -40; e <= -5; ++e) {
    const float x = std::ldexp(1.f, e) * 1.37f;
    const double ref = std::log1p(double(x));
    const double got = double(akx::akx_log1p<float>(x));
    worst_series = std::fmax(worst_series, std::fabs(got - ref) / ref / FLT_EPSILON);
    ++n;
  }
  std::printf("%.3f %.3f %ld\n", worst_series, worst_lib, n);
  return 0;
}
