// mc_math.h — JavaScript-faithful FP64 math for host and gfx950 device code.
//
// The reference evaluates its physics with V8's Math.sin / Math.cos / Math.tan /
// Math.pow / Math.hypot, which are ports of Sun's fdlibm 5.3 (V8 src/base/ieee754.cc).
// Per-cell chatter verdicts near the 0.45 threshold (src/stabtd.js:10) are only robust
// when the GPU reproduces those doubles bit for bit, so this header restates the
// published fdlibm method: argument reduction by pi/2 (e_rem_pio2.c, medium range),
// the minimax kernels k_sin.c / k_cos.c / k_tan.c and e_pow.c, plus V8's hypot
// (scale by the largest magnitude, compensated sum, sqrt, rescale) and JavaScript's
// double % (an exact remainder, here by exact power-of-two multiples of the divisor).
//
// Everything is written so that `hipcc -ffp-contract=off` produces the same doubles on
// the host and on the device: no fused multiply-add, IEEE division and sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MC_HD __host__ __device__ __forceinline__

namespace mcm {

constexpr double JS_PI = 3.141592653589793;
constexpr double TWO_PI = 2 * JS_PI;

MC_HD int32_t hiw(double x) { return (int32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }
MC_HD uint32_t low(double x) { return (uint32_t)__builtin_bit_cast(uint64_t, x); }
MC_HD double words(uint32_t h, uint32_t l) {
  return __builtin_bit_cast(double, ((uint64_t)h << 32) | (uint64_t)l);
}
MC_HD double clear_lo(double x) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) & 0xffffffff00000000ull);
}
MC_HD double fabs_(double x) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) & 0x7fffffffffffffffull);
}

// ---- reduction by pi/2 for |x| <= 2^19 * pi/2 --------------------------------------
__host__ __device__ inline int rem_pio2(double x, double& y0, double& y1) {
  constexpr double invpio2 = 6.36619772367581382433e-01;  // 0x3FE45F30 6DC9C883
  constexpr double pio2_1 = 1.57079632673412561417e+00;   // 0x3FF921FB 54400000
  constexpr double pio2_1t = 6.07710050650619224932e-11;  // 0x3DD0B461 1A626331
  constexpr double pio2_2 = 6.07710050630396597660e-11;   // 0x3DD0B461 1A600000
  constexpr double pio2_2t = 2.02226624879595063154e-21;  // 0x3BA3198A 2E037073
  constexpr double pio2_3 = 2.02226624871116645580e-21;   // 0x3BA3198A 2E000000
  constexpr double pio2_3t = 8.47842766036889956997e-32;  // 0x397B839A 252049C1
  const int32_t hx = hiw(x), ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) { y0 = x; y1 = 0; return 0; }
  if (ix < 0x4002d97c) {  // |x| < 3pi/4: n = +-1
    double z;
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) { y0 = z - pio2_1t; y1 = (z - y0) - pio2_1t; }
      else { z -= pio2_2; y0 = z - pio2_2t; y1 = (z - y0) - pio2_2t; }
      return 1;
    }
    z = x + pio2_1;
    if (ix != 0x3ff921fb) { y0 = z + pio2_1t; y1 = (z - y0) + pio2_1t; }
    else { z += pio2_2; y0 = z + pio2_2t; y1 = (z - y0) + pio2_2t; }
    return -1;
  }
  // medium size (the only other range the engine produces)
  double t = fabs_(x);
  const int n = (int)(t * invpio2 + 0.5);
  const double fn = (double)n;
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t;
  // fdlibm's table npio2_hw[n-1] holds the high words of n*pi/2 (n = 1..32), which
  // equal hiw(n * fl(pi/2)) for every n; computing them keeps the device code free of
  // memory loads on this path.  The cancellation check runs only near those multiples.
  if (n < 32 && ix != hiw((double)n * 1.5707963267948966)) {
    y0 = r - w;
  } else {
    const int j = ix >> 20;
    y0 = r - w;
    int i = j - ((hiw(y0) >> 20) & 0x7ff);
    if (i > 16) {
      t = r; w = fn * pio2_2; r = t - w;
      w = fn * pio2_2t - ((t - r) - w);
      y0 = r - w;
      i = j - ((hiw(y0) >> 20) & 0x7ff);
      if (i > 49) {
        t = r; w = fn * pio2_3; r = t - w;
        w = fn * pio2_3t - ((t - r) - w);
        y0 = r - w;
      }
    }
  }
  y1 = (r - y0) - w;
  if (hx < 0) { y0 = -y0; y1 = -y1; return -n; }
  return n;
}

MC_HD double k_sin(double x, double y, int iy) {
  constexpr double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                   S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                   S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix < 0x3e400000 && (int)x == 0) return x;
  const double z = x * x, v = z * x;
  const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

MC_HD double k_cos(double x, double y) {
  constexpr double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                   C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                   C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix < 0x3e400000 && (int)x == 0) return 1.0;
  const double z = x * x;
  const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  const double qx = ix > 0x3fe90000 ? 0.28125 : words((uint32_t)(ix - 0x00200000), 0);
  const double hz = 0.5 * z - qx, a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

// sin and cos of one argument, sharing the reduction (each equals Math.sin / Math.cos)
MC_HD void sincos_js(double x, double& s, double& c) {
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) { s = k_sin(x, 0.0, 0); c = k_cos(x, 0.0); return; }
  if (ix >= 0x7ff00000) { s = c = x - x; return; }
  double y0, y1;
  const int n = rem_pio2(x, y0, y1);
  const double ks = k_sin(y0, y1, 1), kc = k_cos(y0, y1);
  switch (n & 3) {
    case 0: s = ks; c = kc; break;
    case 1: s = kc; c = -ks; break;
    case 2: s = -ks; c = -kc; break;
    default: s = -kc; c = ks; break;
  }
}

MC_HD double cos_js(double x) { double s, c; sincos_js(x, s, c); return c; }
MC_HD double sin_js(double x) { double s, c; sincos_js(x, s, c); return s; }

__host__ __device__ inline double k_tan(double x, double y, int iy) {
  constexpr double T[13] = {
      3.33333333333334091986e-01, 1.33333333333201242699e-01, 5.39682539762260521377e-02,
      2.18694882948595424599e-02, 8.86323982359930005737e-03, 3.59207910759131235356e-03,
      1.45620945432529025516e-03, 5.88041240820264096874e-04, 2.46463134818469906812e-04,
      7.81794442939557092300e-05, 7.14072491382608190305e-05, -1.85586374855275456654e-05,
      2.59073051863633712884e-05};
  constexpr double pio4 = 7.85398163397448278999e-01, pio4lo = 3.06161699786838301793e-17;
  double z, r, v, w, s;
  const int32_t hx = hiw(x), ix = hx & 0x7fffffff;
  if (ix < 0x3e300000 && (int)x == 0) {
    if (((ix | (int32_t)low(x)) | (iy + 1)) == 0) return 1.0 / fabs_(x);
    if (iy == 1) return x;
    z = w = x + y;
    z = clear_lo(z);
    v = y - (z - x);
    double t, a;
    t = a = -1.0 / w;
    t = clear_lo(t);
    s = 1.0 + t * z;
    return t + a * (s + t * v);
  }
  if (ix >= 0x3FE59428) {
    if (hx < 0) { x = -x; y = -y; }
    z = pio4 - x;
    w = pio4lo - y;
    x = z + w;
    y = 0.0;
  }
  z = x * x;
  w = z * z;
  r = T[1] + w * (T[3] + w * (T[5] + w * (T[7] + w * (T[9] + w * T[11]))));
  v = z * (T[2] + w * (T[4] + w * (T[6] + w * (T[8] + w * (T[10] + w * T[12])))));
  s = z * x;
  r = y + z * (s * (r + v) + y);
  r += T[0] * s;
  w = x + r;
  if (ix >= 0x3FE59428) {
    v = (double)iy;
    return (double)(1 - ((hx >> 30) & 2)) * (v - 2.0 * (x - (w * w / (w + v) - r)));
  }
  if (iy == 1) return w;
  z = clear_lo(w);
  v = r - (z - x);
  double t, a;
  t = a = -1.0 / w;
  t = clear_lo(t);
  s = 1.0 + t * z;
  return t + a * (s + t * v);
}

__host__ __device__ inline double tan_js(double x) {
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return k_tan(x, 0.0, 1);
  if (ix >= 0x7ff00000) return x - x;
  double y0, y1;
  const int n = rem_pio2(x, y0, y1);
  return k_tan(y0, y1, 1 - ((n & 1) << 1));
}

// pow(x, y) for the size-effect law (h0/h)^mc: x > 0 finite, y finite (e_pow.c; the
// special-value prologue keeps only the branches reachable with such arguments)
__host__ __device__ inline double pow_js(double x, double y) {
  // bp = {1, 1.5}, dp_h = {0, log2(1.5) high}, dp_l = {0, its low part}; selected by
  // value rather than indexed, so no table load reaches the device code
  constexpr double dp_h1 = 5.84962487220764160156e-01, dp_l1 = 1.35003920212974897128e-08;
  constexpr double two53 = 9007199254740992.0;
  constexpr double L1 = 5.99999999999994648725e-01, L2 = 4.28571428578550184252e-01,
                   L3 = 3.33333329818377432918e-01, L4 = 2.72728123808534006489e-01,
                   L5 = 2.30660745775561754067e-01, L6 = 2.06975017800338417784e-01;
  constexpr double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                   P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                   P5 = 4.13813679705723846039e-08;
  constexpr double lg2 = 6.93147180559945286227e-01, lg2_h = 6.93147182464599609375e-01,
                   lg2_l = -1.90465429995776804525e-09;
  constexpr double ovt = 8.0085662595372944372e-17;
  constexpr double cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01,
                   cp_l = -7.02846165095275826516e-09;
  constexpr double ivln2 = 1.44269504088896338700e+00, ivln2_h = 1.44269502162933349609e+00,
                   ivln2_l = 1.92596299112661746887e-08;
  constexpr double huge = 1.0e300, tiny = 1.0e-300;
  double z, ax, z_h, z_l, p_h, p_l, y1, t1, t2, r, t, u, v, w;
  int32_t i, j, k, n;
  const int32_t hx = hiw(x), hy = hiw(y);
  const uint32_t lx = low(x), ly = low(y);
  int32_t ix = hx & 0x7fffffff;
  const int32_t iy = hy & 0x7fffffff;
  if ((iy | (int32_t)ly) == 0) return 1.0;
  if (ix > 0x7ff00000 || (ix == 0x7ff00000 && lx != 0) || iy > 0x7ff00000 ||
      (iy == 0x7ff00000 && ly != 0))
    return x + y;
  if (ly == 0) {
    if (iy == 0x3ff00000) return hy < 0 ? 1.0 / x : x;
    if (hy == 0x40000000) return x * x;
    if (hy == 0x3fe00000 && hx >= 0) return __builtin_sqrt(x);
  }
  ax = fabs_(x);
  if (lx == 0 && (ix == 0x7ff00000 || ix == 0 || ix == 0x3ff00000)) {
    z = ax;
    if (hy < 0) z = 1.0 / z;
    return z;  // x > 0 here
  }
  if (iy > 0x41e00000) {  // |y| > 2^31
    if (iy > 0x43f00000) {
      if (ix <= 0x3fefffff) return hy < 0 ? huge * huge : tiny * tiny;
      if (ix >= 0x3ff00000) return hy > 0 ? huge * huge : tiny * tiny;
    }
    if (ix < 0x3fefffff) return hy < 0 ? huge * huge : tiny * tiny;
    if (ix > 0x3ff00000) return hy > 0 ? huge * huge : tiny * tiny;
    t = ax - 1.0;
    w = (t * t) * (0.5 - t * (0.3333333333333333333333 - t * 0.25));
    u = ivln2_h * t;
    v = t * ivln2_l - w * ivln2;
    t1 = clear_lo(u + v);
    t2 = v - (t1 - u);
  } else {
    double ss, s2, s_h, s_l, t_h, t_l;
    n = 0;
    if (ix < 0x00100000) { ax *= two53; n -= 53; ix = hiw(ax); }
    n += (ix >> 20) - 0x3ff;
    j = ix & 0x000fffff;
    ix = j | 0x3ff00000;
    if (j <= 0x3988E) k = 0;
    else if (j < 0xBB67A) k = 1;
    else { k = 0; n += 1; ix -= 0x00100000; }
    ax = words((uint32_t)ix, low(ax));
    const double bpk = k ? 1.5 : 1.0, dphk = k ? dp_h1 : 0.0, dplk = k ? dp_l1 : 0.0;
    u = ax - bpk;
    v = 1.0 / (ax + bpk);
    ss = u * v;
    s_h = clear_lo(ss);
    t_h = words((uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)), 0);
    t_l = ax - (t_h - bpk);
    s_l = v * ((u - s_h * t_h) - s_h * t_l);
    s2 = ss * ss;
    r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
    r += s_l * (s_h + ss);
    s2 = s_h * s_h;
    t_h = clear_lo(3.0 + s2 + r);
    t_l = r - ((t_h - 3.0) - s2);
    u = s_h * t_h;
    v = s_l * t_h + t_l * ss;
    p_h = clear_lo(u + v);
    p_l = v - (p_h - u);
    z_h = cp_h * p_h;
    z_l = cp_l * p_h + p_l * cp + dplk;
    t = (double)n;
    t1 = clear_lo(((z_h + z_l) + dphk) + t);
    t2 = z_l - (((t1 - t) - dphk) - z_h);
  }
  y1 = clear_lo(y);
  p_l = (y - y1) * t1 + y * t2;
  p_h = y1 * t1;
  z = p_l + p_h;
  j = hiw(z);
  i = (int32_t)low(z);
  if (j >= 0x40900000) {
    if (((j - 0x40900000) | i) != 0) return huge * huge;
    if (p_l + ovt > z - p_h) return huge * huge;
  } else if ((j & 0x7fffffff) >= 0x4090cc00) {
    if (((j - (int32_t)0xc090cc00) | i) != 0) return tiny * tiny;
    if (p_l <= z - p_h) return tiny * tiny;
  }
  i = j & 0x7fffffff;
  k = (i >> 20) - 0x3ff;
  n = 0;
  if (i > 0x3fe00000) {
    n = j + (0x00100000 >> (k + 1));
    k = ((n & 0x7fffffff) >> 20) - 0x3ff;
    t = words((uint32_t)(n & ~(0x000fffff >> k)), 0);
    n = ((n & 0x000fffff) | 0x00100000) >> (20 - k);
    if (j < 0) n = -n;
    p_h -= t;
  }
  t = clear_lo(p_l + p_h);
  u = t * lg2_h;
  v = (p_l - (t - p_h)) * lg2 + t * lg2_l;
  z = u + v;
  w = v - (z - u);
  t = z * z;
  t1 = z - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  r = (z * t1) / (t1 - 2.0) - (w + z * w);
  z = 1.0 - (r - z);
  j = hiw(z) + (int32_t)((uint32_t)n << 20);   // n may be negative: shift its bits
  if ((j >> 20) <= 0) return __builtin_ldexp(z, n);  // subnormal result
  return words((uint32_t)j, low(z));
}

// pow_js for the operands of the chip-thickness size effect, (H_REF / h)^mc: x a positive
// normal number in [2^-96, 2^96] and y with 0 < |y| <= 8 that is none of -1, 1, 2, 1/2
// (pow_cell_ok).  There |y * log2 x| < 1024, so the result is a normal number and none of
// pow_js's special cases, overflow or underflow branches is taken; the general path is
// kept operation for operation with its remaining branches turned into selects, so the
// result equals pow_js(x, y) bit for bit.  y1 is clear_lo(y) and yr is y - y1.
MC_HD bool pow_cell_ok(double y) {
  if (!(y == y) || y == 0 || fabs_(y) > 8) return false;
  return !(y == 1 || y == -1 || y == 2 || y == 0.5);
}
MC_HD double pow_cell(double x, double y, double y1, double yr) {
  constexpr double dp_h1 = 5.84962487220764160156e-01, dp_l1 = 1.35003920212974897128e-08;
  constexpr double L1 = 5.99999999999994648725e-01, L2 = 4.28571428578550184252e-01,
                   L3 = 3.33333329818377432918e-01, L4 = 2.72728123808534006489e-01,
                   L5 = 2.30660745775561754067e-01, L6 = 2.06975017800338417784e-01;
  constexpr double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                   P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                   P5 = 4.13813679705723846039e-08;
  constexpr double lg2 = 6.93147180559945286227e-01, lg2_h = 6.93147182464599609375e-01,
                   lg2_l = -1.90465429995776804525e-09;
  constexpr double cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01,
                   cp_l = -7.02846165095275826516e-09;
  // log2(x) = t1 + t2 (x normal, so no subnormal scaling)
  int32_t ix = hiw(x);
  int32_t n = (ix >> 20) - 0x3ff;
  const int32_t jm = ix & 0x000fffff;
  const bool wrap = jm >= 0xBB67A;            // mantissa above sqrt(3): use 2^(n+1)
  const bool k1 = jm > 0x3988E && !wrap;      // mantissa in (sqrt(3/2), sqrt(3)): base 1.5
  n += wrap ? 1 : 0;
  ix = (jm | 0x3ff00000) - (wrap ? 0x00100000 : 0);
  const double ax = words((uint32_t)ix, low(x));
  const double bpk = k1 ? 1.5 : 1.0, dphk = k1 ? dp_h1 : 0.0, dplk = k1 ? dp_l1 : 0.0;
  double u = ax - bpk;
  double v = 1.0 / (ax + bpk);
  const double ss = u * v;
  const double s_h = clear_lo(ss);
  double t_h = words((uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k1 ? (1 << 18) : 0)), 0);
  double t_l = ax - (t_h - bpk);
  const double s_l = v * ((u - s_h * t_h) - s_h * t_l);
  double s2 = ss * ss;
  double r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
  r += s_l * (s_h + ss);
  s2 = s_h * s_h;
  t_h = clear_lo(3.0 + s2 + r);
  t_l = r - ((t_h - 3.0) - s2);
  u = s_h * t_h;
  v = s_l * t_h + t_l * ss;
  double p_h = clear_lo(u + v);
  double p_l = v - (p_h - u);
  const double z_h = cp_h * p_h;
  const double z_l = cp_l * p_h + p_l * cp + dplk;
  double t = (double)n;
  const double t1 = clear_lo(((z_h + z_l) + dphk) + t);
  const double t2 = z_l - (((t1 - t) - dphk) - z_h);
  // y * log2(x) = p_h + p_l
  p_l = yr * t1 + y * t2;
  p_h = y1 * t1;
  double z = p_l + p_h;
  const int32_t j = hiw(z);
  const int32_t i = j & 0x7fffffff;
  // |z| > 1/2: split off the integer part n (|z| < 1024, so every shift is in range)
  const bool big = i > 0x3fe00000;
  const int32_t kz = big ? (i >> 20) - 0x3ff : -1;
  int32_t nn = j + (0x00100000 >> (kz + 1));
  const int32_t kn = big ? ((nn & 0x7fffffff) >> 20) - 0x3ff : 0;
  t = words((uint32_t)(nn & ~(0x000fffff >> kn)), 0);
  nn = ((nn & 0x000fffff) | 0x00100000) >> (20 - kn);
  if (j < 0) nn = -nn;
  n = big ? nn : 0;
  p_h = big ? p_h - t : p_h;
  // 2^z
  t = clear_lo(p_l + p_h);
  u = t * lg2_h;
  v = (p_l - (t - p_h)) * lg2 + t * lg2_l;
  z = u + v;
  const double w = v - (z - u);
  t = z * z;
  const double e1 = z - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  r = (z * e1) / (e1 - 2.0) - (w + z * w);
  z = 1.0 - (r - z);
  const double res = words((uint32_t)hiw(z) + ((uint32_t)n << 20), low(z));
  return x == 1.0 ? 1.0 : res;
}

// Math.atan (fdlibm s_atan: reduction to |x| < 7/16 around the break points 1/2, 1, 3/2
// and infinity, then an odd polynomial in x^2 split into even and odd halves)
MC_HD double atan_js(double x) {
  constexpr double hi0 = 4.63647609000806093515e-01, hi1 = 7.85398163397448278999e-01,
                   hi2 = 9.82793723247329054082e-01, hi3 = 1.57079632679489655800e+00;
  constexpr double lo0 = 2.26987774529616870924e-17, lo1 = 3.06161699786838301793e-17,
                   lo2 = 1.39033110312309984516e-17, lo3 = 6.12323399573676603587e-17;
  constexpr double T0 = 3.33333333333329318027e-01, T1 = -1.99999999998764832476e-01,
                   T2 = 1.42857142725034663711e-01, T3 = -1.11111104054623557880e-01,
                   T4 = 9.09088713343650656196e-02, T5 = -7.69187620504482999495e-02,
                   T6 = 6.66107313738753120669e-02, T7 = -5.83357013379057348645e-02,
                   T8 = 4.97687799461593236017e-02, T9 = -3.65315727442169155270e-02,
                   T10 = 1.62858201153657823623e-02;
  const int32_t hx = hiw(x), ix = hx & 0x7fffffff;
  if (ix >= 0x44100000) {  // |x| >= 2^66
    if (ix > 0x7ff00000 || (ix == 0x7ff00000 && low(x) != 0)) return x + x;
    return hx > 0 ? hi3 + lo3 : -hi3 - lo3;
  }
  int id;
  if (ix < 0x3fdc0000) {   // |x| < 7/16
    if (ix < 0x3e200000) return x;
    id = -1;
  } else {
    x = fabs_(x);
    if (ix < 0x3ff30000) {
      if (ix < 0x3fe60000) { id = 0; x = (2.0 * x - 1.0) / (2.0 + x); }
      else { id = 1; x = (x - 1.0) / (x + 1.0); }
    } else if (ix < 0x40038000) { id = 2; x = (x - 1.5) / (1.0 + 1.5 * x); }
    else { id = 3; x = -1.0 / x; }
  }
  const double z = x * x, w = z * z;
  const double s1 = z * (T0 + w * (T2 + w * (T4 + w * (T6 + w * (T8 + w * T10)))));
  const double s2 = w * (T1 + w * (T3 + w * (T5 + w * (T7 + w * T9))));
  if (id < 0) return x - x * (s1 + s2);
  const double ah = id == 0 ? hi0 : id == 1 ? hi1 : id == 2 ? hi2 : hi3;
  const double al = id == 0 ? lo0 : id == 1 ? lo1 : id == 2 ? lo2 : lo3;
  const double r = ah - ((x * (s1 + s2) - al) - x);
  return hx < 0 ? -r : r;
}

// Math.hypot(a, b) as V8's builtin computes it
MC_HD double hypot_js(double a, double b) {
  const double fa = fabs_(a), fb = fabs_(b);
  if (fa == __builtin_inf() || fb == __builtin_inf()) return __builtin_inf();
  if (a != a || b != b) return __builtin_nan("");
  const double mx = fa > fb ? fa : fb;
  if (mx == 0) return 0;
  double sum = 0, comp = 0, nrm, summand, prelim;
  nrm = fa / mx; summand = nrm * nrm - comp; prelim = sum + summand;
  comp = (prelim - sum) - summand; sum = prelim;
  nrm = fb / mx; summand = nrm * nrm - comp; prelim = sum + summand;
  sum = prelim;
  return __builtin_sqrt(sum) * mx;
}

// JavaScript x % (2*pi) for the engine's angles: an exact remainder, obtained by
// subtracting exact power-of-two multiples of 2*pi (each subtraction is exact by
// Sterbenz's lemma because the running value lies in [m, 2m) for the multiple m)
MC_HD double fmod_2pi(double x) {
  double r = fabs_(x);
  if (r >= TWO_PI) {
    double m = TWO_PI;
    while (m * 2 <= r) m *= 2;
    while (m >= TWO_PI) {
      if (r >= m) r -= m;
      m *= 0.5;
    }
  }
  return __builtin_copysign(r, x);
}

// Math.atan2 (fdlibm e_atan2 as FreeBSD's msun keeps it, which V8 ports: the quadrant from
// the signs, the special values, then atan(|y/x|) corrected by pi with its low part)
MC_HD double atan2_js(double y, double x) {
  constexpr double tiny = 1.0e-300, pi_o_4 = 7.8539816339744827900e-01, pi_o_2 = 1.5707963267948965580e+00,
                   pi = 3.1415926535897931160e+00, pi_lo = 1.2246467991473531772e-16;
  const int32_t hx = hiw(x), ix = hx & 0x7fffffff, hy = hiw(y), iy = hy & 0x7fffffff;
  const uint32_t lx = low(x), ly = low(y);
  if ((uint32_t)(ix | ((lx | (0u - lx)) >> 31)) > 0x7ff00000u || (uint32_t)(iy | ((ly | (0u - ly)) >> 31)) > 0x7ff00000u)
    return x + y;   // a NaN
  if ((((uint32_t)hx - 0x3ff00000u) | lx) == 0) return atan_js(y);   // x = 1
  int m = ((hy >> 31) & 1) | ((hx >> 30) & 2);                      // 2 sign(x) + sign(y)
  if ((iy | (int32_t)ly) == 0) {                                     // y = 0
    if (m < 2) return y;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if ((ix | (int32_t)lx) == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;   // x = 0
  if (ix == 0x7ff00000) {                                            // x infinite
    if (iy == 0x7ff00000) {
      if (m == 0) return pi_o_4 + tiny;
      if (m == 1) return -pi_o_4 - tiny;
      if (m == 2) return 3.0 * pi_o_4 + tiny;
      return -3.0 * pi_o_4 - tiny;
    }
    if (m == 0) return 0.0;
    if (m == 1) return -0.0;
    if (m == 2) return pi + tiny;
    return -pi - tiny;
  }
  if (iy == 0x7ff00000) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;   // y infinite
  const int k = (iy - ix) >> 20;
  double z;
  if (k > 60) { z = pi_o_2 + 0.5 * pi_lo; m &= 1; }   // |y/x| > 2^60: +-pi/2 whatever x's sign
  else if (hx < 0 && k < -60) z = 0.0;           // |y|/x < -2^60
  else z = atan_js(fabs_(y / x));
  if (m == 0) return z;
  if (m == 1) return -z;
  if (m == 2) return pi - (z - pi_lo);
  return (z - pi_lo) - pi;
}

// Math.acos (fdlibm e_acos: a rational approximation of asin around 0 for |x| < 1/2, and
// through sqrt((1 -+ x) / 2) beyond, with pi/2 split in high and low parts)
MC_HD double acos_js(double x) {
  constexpr double one = 1.0, pi = 3.14159265358979311600e+00, pio2_hi = 1.57079632679489655800e+00,
                   pio2_lo = 6.12323399573676603587e-17;
  constexpr double pS0 = 1.66666666666666657415e-01, pS1 = -3.25565818622400915405e-01,
                   pS2 = 2.01212532134862925881e-01, pS3 = -4.00555345006794114027e-02,
                   pS4 = 7.91534994289814532176e-04, pS5 = 3.47933107596021167570e-05;
  constexpr double qS1 = -2.40339491173441421878e+00, qS2 = 2.02094576023350569471e+00,
                   qS3 = -6.88283971605453293030e-01, qS4 = 7.70381505559019352791e-02;
  const int32_t hx = hiw(x), ix = hx & 0x7fffffff;
  if (ix >= 0x3ff00000) {                                            // |x| >= 1
    if (((ix - 0x3ff00000) | (int32_t)low(x)) == 0) return hx > 0 ? 0.0 : pi + 2.0 * pio2_lo;
    return (x - x) / (x - x);                                        // NaN
  }
  if (ix < 0x3fe00000) {                                             // |x| < 1/2
    if (ix <= 0x3c600000) return pio2_hi + pio2_lo;
    const double z = x * x;
    const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    const double q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    const double r = p / q;
    return pio2_hi - (x - (pio2_lo - x * r));
  }
  if (hx < 0) {                                                      // x < -1/2
    const double z = (one + x) * 0.5;
    const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    const double q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    const double s = __builtin_sqrt(z);
    const double r = p / q;
    const double w = r * s - pio2_lo;
    return pi - 2.0 * (s + w);
  }
  const double z = (one - x) * 0.5;                                  // x > 1/2
  const double s = __builtin_sqrt(z);
  const double df = clear_lo(s);
  const double c = (z - df * df) / (s + df);
  const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
  const double q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
  const double r = p / q;
  const double w = r * s + c;
  return 2.0 * (df + w);
}

// Math.round for the engine's positive arguments: halves round up
MC_HD double round_js(double x) {
  double r = __builtin_floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

}  // namespace mcm
