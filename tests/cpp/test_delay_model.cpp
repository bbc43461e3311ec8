// DelayModel::calculateDelayInterpolator (model.cpp:229-304) against what the
// mathematics fixes: an order-2 interpolator reproduces a quadratic delay
// polynomial exactly at every increment, order 1 a linear one, order 0 is the
// midpoint value; the model interval is chosen by the time; times outside the
// model and orders other than 0..2 are refused.  The interval index is the
// time over the interval length truncated toward zero (model.cpp:244), so a
// time less than one interval before the start still maps to the first.
#include <cmath>
#include <cstdio>

#include "gpumode.h"

using namespace fxgpu;

static int fails = 0;
static void expect(bool ok, const char* what) {
  printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) fails++;
}
static bool close(double a, double b, double tol) { return std::fabs(a - b) <= tol * (1.0 + std::fabs(b)); }

int main() {
  // a quadratic delay polynomial in microseconds over 120 s intervals
  const double q0 = 1234.56789, q1 = -0.731, q2 = 2.5e-6;
  auto quad = [&](double t) { return q0 + q1 * t + q2 * t * t; };
  DelayModel m({{q0, q1, q2}, {q0 + 120 * q1 + 14400 * q2, q1 + 240 * q2, q2}}, 120.0);
  double c[3];
  const double t0 = 17.25, span = 0.128;
  const int n = 2000;
  expect(m.calculateDelayInterpolator(t0, span, n, 2, c), "order 2 inside the first interval");
  bool ok = true;
  for (int i = 0; i <= n; i += 37) {
    double want = quad(t0 + span * i / n);
    double got = c[0] * i * i + c[1] * i + c[2];
    ok &= close(got, want, 1e-12);
  }
  expect(ok, "order 2 reproduces a quadratic delay at every increment");
  // the second interval continues the same polynomial: same values there
  expect(m.calculateDelayInterpolator(130.0, span, n, 2, c), "order 2 inside the second interval");
  ok = true;
  for (int i = 0; i <= n; i += 101) ok &= close(c[0] * i * i + c[1] * i + c[2], quad(130.0 + span * i / n), 1e-12);
  expect(ok, "second interval selected by time");
  expect(m.calculateDelayInterpolator(240.0, 0.0, n, 0, c), "exactly at the end of the last interval");
  expect(close(c[0], quad(240.0), 1e-12), "end value");
  expect(!m.calculateDelayInterpolator(241.0, span, n, 2, c), "after the model: refused");
  expect(m.calculateDelayInterpolator(-1.0, span, n, 2, c) && close(c[2], quad(-1.0), 1e-12),
         "less than one interval before the model: first interval, extrapolated");
  expect(!m.calculateDelayInterpolator(-121.0, span, n, 2, c), "a whole interval before the model: refused");
  expect(!m.calculateDelayInterpolator(t0, span, n, 3, c), "order 3 refused");

  // order 1 on a linear delay: exact line
  DelayModel lin({{50.0, 0.25}}, 100.0);
  expect(lin.calculateDelayInterpolator(3.0, 0.5, 64, 1, c), "order 1");
  ok = true;
  for (int i = 0; i <= 64; i++) ok &= close(c[0] * i + c[1], 50.0 + 0.25 * (3.0 + 0.5 * i / 64), 1e-13);
  expect(ok, "order 1 reproduces a linear delay");
  // order 1 on a quadratic: slope is the chord, offset shifted by a third of the sag
  expect(m.calculateDelayInterpolator(t0, span, n, 1, c), "order 1 on a quadratic");
  double s0 = quad(t0), s1 = quad(t0 + span / 2), s2 = quad(t0 + span);
  double sag = s1 - (s0 + s2) / 2;
  expect(close(c[0] * n, s2 - s0, 1e-12) && close(c[1], s0 + sag / 3, 1e-12), "order 1 chord and sag/3 offset");
  // order 0: the midpoint
  expect(m.calculateDelayInterpolator(t0, span, n, 0, c) && close(c[0], quad(t0 + span / 2), 1e-13), "order 0 midpoint");
  // a cubic: order 2 matches at start, middle and end (three-point interpolation)
  DelayModel cub({{1.0, 2.0, 3.0, 4.0}}, 10.0);
  auto cubic = [](double t) { return 1.0 + 2.0 * t + 3.0 * t * t + 4.0 * t * t * t; };
  expect(cub.calculateDelayInterpolator(1.5, 2.0, 10, 2, c), "order 2 on a cubic");
  expect(close(c[2], cubic(1.5), 1e-13) && close(25 * c[0] + 5 * c[1] + c[2], cubic(2.5), 1e-13) &&
             close(100 * c[0] + 10 * c[1] + c[2], cubic(3.5), 1e-13),
         "order 2 matches a cubic at the three sample points");
  printf("%d failure(s)\n", fails);
  return fails ? 1 : 0;
}
