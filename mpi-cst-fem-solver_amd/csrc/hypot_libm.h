// hypot_libm.h -- sqrt(x^2 + y^2) as the host C library of this toolchain computes it, for host
// and device code.
//
// The traction quadrature takes each edge's length from std::hypot (bc.cpp:231), so the loads
// equal the reference's bit for bit only if the device returns the same double.  That C
// library (glibc 2.35, x86-64 without FMA) is not correctly rounded; it implements the
// correction of C. F. Borges, "An Improved Algorithm for hypot(a,b)" (arXiv:1904.09481,
// 2019, the FMA-free correction of its Section 4):
//   h = sqrt(a^2 + b^2) for a >= b, then h -= (t1 + t2) / (2h), where t1 + t2 approximates
//   h^2 - a^2 - b^2 from the exact-in-spirit splittings below (two branches on h <= 2b),
// with power-of-two scaling by 2^600 when a > 2^511 (squares overflow) or b < 2^-459 (the
// correction's d^2 ~ h^2 2^-104 would underflow), and a + b returned when b <= 2^-54 a.  Every operation must be rounded separately (build with -ffp-contract=off):
// a fused multiply-add changes the result.  tests/test_hypot_libm.py checks the host build
// against the C library.  Define CSTFEM_HD before inclusion to qualify it for device code.
#pragma once

#include <cmath>

#ifndef CSTFEM_HD
#define CSTFEM_HD
#endif

namespace hypot_detail {

// a >= b >= 0, squares of a, b and a - b representable
CSTFEM_HD inline double corrected(double a, double b) {
  double h = std::sqrt(a * a + b * b);
  double t1, t2;
  if (h <= 2.0 * b) {
    const double d = h - b;
    t1 = a * (2.0 * d - a);
    t2 = (d - 2.0 * (a - b)) * d;
  } else {
    const double d = h - a;
    t1 = 2.0 * d * (a - 2.0 * b);
    t2 = (4.0 * d - b) * b + d * d;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h;
}

}  // namespace hypot_detail

CSTFEM_HD inline double hypot_libm(double x, double y) {
  if (!std::isfinite(x) || !std::isfinite(y)) {
    if (std::isinf(x) || std::isinf(y)) return INFINITY;
    return x + y;
  }
  x = std::fabs(x);
  y = std::fabs(y);
  const double a = x < y ? y : x, b = x < y ? x : y;
  const double eps = 0x1p-54, scale = 0x1p-600;
  if (a > 0x1p+511) {
    if (b <= a * eps) return a + b;
    return hypot_detail::corrected(a * scale, b * scale) / scale;
  }
  if (b < 0x1p-459) {
    if (a >= b / eps) return a + b;
    return hypot_detail::corrected(a / scale, b / scale) * scale;
  }
  if (b <= a * eps) return a + b;
  return hypot_detail::corrected(a, b);
}
