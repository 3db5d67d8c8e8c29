"""The column code's device logarithm (ln_fast in gitm_amd/csrc/vert_column.h)
on the CPU: the same source built with g++ (its host build differs only in
the reciprocal seed, a float reciprocal in place of the device's, which two
Newton steps refine to full precision either way), against long-double logl
on 6e6 arguments over the whole double range, near 1, and over the range of
the step's densities and temperatures.  The bar is below one ulp
everywhere (0.83 measured); a wrong split of ln 2 or a reduction to [1/2, 1)
instead of [sqrt(1/2), sqrt(2)) fails it by far, and the check of the
checker below moves the result by two ulps.
The special values follow log: +0 and -0 give -inf, negatives and NaN give
NaN, +inf gives +inf; denormals and the largest double are exact to an
ulp."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#define GV_HOST_EMULATION 1
#include "vert_column.h"

static double ulp_err(double got, long double ref) {
  const double r = (double)ref;
  const double u = nextafter(fabs(r), INFINITY) - fabs(r);
  return (double)(fabsl((long double)got - ref) / u);
}

int main() {
  srand48(7);
  double worst = 0.0, at = 0.0;
  for (int range = 0; range < 4; range++)
    for (long i = 0; i < 1500000; i++) {
      const double u = drand48();
      double x;
      if (range == 0) x = exp((u - 0.5) * 1480.0);         // 1e-321 .. 1e321
      else if (range == 1) x = 0.5 + u;                    // around 1
      else if (range == 2) x = 1.0 + (u - 0.5) * 1e-6;     // next to 1
      else x = exp(u * 70.0 - 5.0);                        // densities, temperatures
      if (!(x > 0.0) || x == INFINITY) continue;
      const double e = ulp_err(PERTURB(gv::ln_fast(x)), logl((long double)x));
      if (e > worst) { worst = e; at = x; }
    }
  printf("%.6f %.17g\n", worst, at);
  const double sp[] = {0.0, -0.0, -1.0, INFINITY, NAN, 1.0, 2.0, 0.5,
                       4.9406564584124654e-324, 2.2250738585072014e-308,
                       1.7976931348623157e308};
  for (double x : sp) printf("%.17g %.17g\n", gv::ln_fast(x), log(x));
  return 0;
}
"""


def run(tmp_path, perturb: str = "(v)") -> tuple[float, list[tuple[float, float]]]:
    src = tmp_path / "ln_fast.cpp"
    exe = tmp_path / "ln_fast"
    src.write_text(f"#define PERTURB(v) {perturb}\n" + PROGRAM)
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "gitm_amd", "csrc"), str(src), "-o", str(exe),
                    "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    worst = float(out[0].split()[0])
    pairs = [tuple(float(t) for t in line.split()) for line in out[1:] if line.strip()]
    return worst, pairs


def same(a: float, b: float) -> bool:
    return a == b or (a != a and b != b)


def test_ln_fast_below_one_ulp(tmp_path):
    worst, pairs = run(tmp_path)
    assert worst < 1.0, worst
    # special values as log; the extremes to an ulp
    for got, want in pairs[:8]:
        assert same(got, want), (got, want)
    for got, want in pairs[8:]:
        assert abs(got - want) <= abs(want) * 2.3e-16, (got, want)


def test_the_bar_catches_an_error_of_two_ulps(tmp_path):
    # the checker itself: the same result moved by two ulps fails the bar
    worst, _ = run(tmp_path, "nextafter(nextafter((v), INFINITY), INFINITY)")
    assert worst > 1.0
