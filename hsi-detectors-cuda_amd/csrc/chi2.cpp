// chi2.cpp - RX threshold from a false-alarm rate (host side of the C ABI).
//
// Under a Gaussian background the RX statistic is chi-squared with `bands`
// degrees of freedom, so the threshold for a per-pixel false-alarm rate q is
// the value t with P(bands/2, t/2) = 1 - q, where P is the regularised lower
// incomplete gamma function. The reference solves this by bisection
// (detectors_cpu.cpp:189-206) over P evaluated by its power series below
// x = a + 1 and by the continued fraction for Q = 1 - P above it, using the
// modified Lentz method (detectors_cpu.cpp:40-73). The same published method
// is used here (Numerical Recipes, 3rd ed., section 6.2) with the same
// iteration limits and tolerances, so the two agree to the last bit.
#include <cmath>
#include <limits>

#include "hsi_mi355x.h"

namespace {

double lower_gamma_regularised(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double prefactor_log = -x + a * std::log(x) - std::lgamma(a);
  if (x < a + 1.0) {
    double term = 1.0 / a;
    double total = term;
    for (int k = 1; k < 1000; ++k) {
      term *= x / (a + k);
      total += term;
      if (std::fabs(term) < std::fabs(total) * 1e-16) break;
    }
    return total * std::exp(prefactor_log);
  }
  constexpr double kFloor = 1e-300;
  double b = x + 1.0 - a;
  double c = 1.0 / kFloor;
  double d = 1.0 / b;
  double h = d;
  for (int k = 1; k < 1000; ++k) {
    const double an = -k * (k - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < kFloor) d = kFloor;
    c = b + an / c;
    if (std::fabs(c) < kFloor) c = kFloor;
    d = 1.0 / d;
    const double step = d * c;
    h *= step;
    if (std::fabs(step - 1.0) < 1e-16) break;
  }
  return 1.0 - std::exp(prefactor_log) * h;
}

}  // namespace

extern "C" double hsi_rx_threshold(int bands, double false_alarm_rate) {
  if (bands <= 0 || !(false_alarm_rate > 0.0) || !(false_alarm_rate < 1.0))
    return std::numeric_limits<double>::quiet_NaN();
  const double a = 0.5 * bands;
  const double target = 1.0 - false_alarm_rate;
  double lo = 0.0;
  double hi = 4.0 * bands > 16.0 ? 4.0 * bands : 16.0;
  while (lower_gamma_regularised(a, 0.5 * hi) < target && hi < 1e9) hi *= 2.0;
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (lo + hi);
    if (lower_gamma_regularised(a, 0.5 * mid) < target) lo = mid;
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}
