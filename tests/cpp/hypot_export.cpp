// Host build of the device path's hypot (csrc/hypot_libm.h), for the CPU test that compares it
// with the C library's hypot.
#include "hypot_libm.h"

extern "C" void hypot_batch(int n, const double* x, const double* y, double* out) {
  for (int i = 0; i < n; ++i) out[i] = hypot_libm(x[i], y[i]);
}
