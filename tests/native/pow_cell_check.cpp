// pow_cell (the search kernel's size-effect power) against pow_js, bit for bit, on the
// host: random operands across its whole domain, chip-thickness-shaped operands, operands
// next to 1.0 and the domain's end points.  Exit status 0 when every result agrees.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <random>
#include "mc_math.h"
using namespace mcm;
int main() {
  std::mt19937_64 rng(7);
  long bad = 0, n = 0;
  double ys[] = {0.1, 0.10, 0.25, 0.3, 0.05, 0.01, 0.7, 1.5, 3.0, 7.9, -0.1, -0.5, -2.0, 0.0999, 0.123456789, -7.5, 0.15, 0.2};
  for (double y : ys) {
    if (!pow_cell_ok(y)) { printf("skip %g\n", y); continue; }
    const double y1 = clear_lo(y), yr = y - y1;
    for (int it = 0; it < 250000; it++) {
      double x;
      int mode = it % 4;
      if (mode == 0) x = std::ldexp(1.0 + (rng() >> 11) * 0x1p-53, (int)(rng() % 192) - 96);
      else if (mode == 1) { double h = 0.002 + (rng() >> 11) * 0x1p-53 * 0.3; x = 0.05 / (h > 0.002 ? h : 0.002); }
      else if (mode == 2) x = 0.05 / std::ldexp(1.0 + (rng() >> 11) * 0x1p-53, -(int)(rng() % 10));
      else { uint64_t b = __builtin_bit_cast(uint64_t, 1.0) + (int64_t)(rng() % 2001) - 1000; x = __builtin_bit_cast(double, b); }
      const double a = pow_js(x, y), b = pow_cell(x, y, y1, yr);
      n++;
      if (__builtin_bit_cast(uint64_t, a) != __builtin_bit_cast(uint64_t, b)) { if (bad < 10) printf("x=%a y=%a js=%a cell=%a\n", x, y, a, b); bad++; }
    }
  }
  double sp[] = {1.0, 25.0, 0.5, 2.0, 0x1p-96, 0x1p96, 0.05/0.1, 0.05/0.15};
  for (double x : sp) for (double y : ys) {
    const double y1 = clear_lo(y), yr = y - y1;
    const double a = pow_js(x, y), b = pow_cell(x, y, y1, yr); n++;
    if (__builtin_bit_cast(uint64_t, a) != __builtin_bit_cast(uint64_t, b)) { printf("special x=%a y=%a js=%a cell=%a\n", x, y, a, b); bad++; }
  }
  printf("checked %ld mismatches %ld\n", n, bad);
  return bad != 0;
}
