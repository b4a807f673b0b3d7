/*
 * oracle/fdlibm.c — TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, the smoke
 * check and bench.py's CPU-baseline leg may load this code; the product never does.
 *
 * JavaScript's Math.sin / Math.cos / Math.tan / Math.pow in V8 (the engine that runs the
 * reference, src/cutting.js:183,197 and src/cutting.js:41) are ports of Sun's fdlibm 5.3
 * (V8 src/base/ieee754.cc).  This file restates the published fdlibm method (argument
 * reduction __ieee754_rem_pio2 for |x| < 2^19*pi/2, the minimax kernels __kernel_sin /
 * __kernel_cos / __kernel_tan and __ieee754_pow) so that the oracle reproduces V8's
 * doubles bit for bit.  Math.hypot follows V8's builtin (scale by the largest
 * magnitude, Kahan-compensated sum, sqrt, rescale).  Pinned by tests/test_oracle_math.py
 * against Node's own Math.* on random and edge inputs.
 *
 * Compile with -ffp-contract=off: JavaScript never fuses multiply-add.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static inline int32_t hi_word(double x) { uint64_t u; memcpy(&u, &x, 8); return (int32_t)(u >> 32); }
static inline uint32_t lo_word(double x) { uint64_t u; memcpy(&u, &x, 8); return (uint32_t)u; }
static inline double from_words(uint32_t hi, uint32_t lo) {
  uint64_t u = ((uint64_t)hi << 32) | lo; double x; memcpy(&x, &u, 8); return x;
}
static inline double with_lo(double x, uint32_t lo) { return from_words((uint32_t)hi_word(x), lo); }
static inline double with_hi(double x, uint32_t hi) { return from_words(hi, lo_word(x)); }

/* ---- argument reduction by pi/2 (fdlibm e_rem_pio2.c, medium range) ---- */
static const int32_t npio2_hw[32] = {
  0x3FF921FB, 0x400921FB, 0x4012D97C, 0x401921FB, 0x401F6A7A, 0x4022D97C,
  0x4025FDBB, 0x402921FB, 0x402C463A, 0x402F6A7A, 0x4031475C, 0x4032D97C,
  0x40346B9C, 0x4035FDBB, 0x40378FDB, 0x403921FB, 0x403AB41B, 0x403C463A,
  0x403DD85A, 0x403F6A7A, 0x40407E4C, 0x4041475C, 0x4042106C, 0x4042D97C,
  0x4043A28C, 0x40446B9C, 0x404534AC, 0x4045FDBB, 0x4046C6CB, 0x40478FDB,
  0x404858EB, 0x404921FB,
};
#define INVPIO2 6.36619772367581382433e-01 /* 0x3FE45F30 6DC9C883 */
#define PIO2_1  1.57079632673412561417e+00 /* 0x3FF921FB 54400000 */
#define PIO2_1T 6.07710050650619224932e-11 /* 0x3DD0B461 1A626331 */
#define PIO2_2  6.07710050630396597660e-11 /* 0x3DD0B461 1A600000 */
#define PIO2_2T 2.02226624879595063154e-21 /* 0x3BA3198A 2E037073 */
#define PIO2_3  2.02226624871116645580e-21 /* 0x3BA3198A 2E000000 */
#define PIO2_3T 8.47842766036889956997e-32 /* 0x397B839A 252049C1 */

/* returns n and y[0]+y[1] = x - n*pi/2; valid for |x| <= 2^19*pi/2 (the engine's
 * angles are in [0, 2*pi)); larger arguments return 0x7fffffff */
int or_rem_pio2(double x, double *y) {
  int32_t hx = hi_word(x), ix = hx & 0x7fffffff;
  double z, w, t, r, fn;
  int n, i, j;
  if (ix <= 0x3fe921fb) { y[0] = x; y[1] = 0; return 0; }
  if (ix < 0x4002d97c) { /* |x| < 3pi/4 */
    if (hx > 0) {
      z = x - PIO2_1;
      if (ix != 0x3ff921fb) { y[0] = z - PIO2_1T; y[1] = (z - y[0]) - PIO2_1T; }
      else { z -= PIO2_2; y[0] = z - PIO2_2T; y[1] = (z - y[0]) - PIO2_2T; }
      return 1;
    } else {
      z = x + PIO2_1;
      if (ix != 0x3ff921fb) { y[0] = z + PIO2_1T; y[1] = (z - y[0]) + PIO2_1T; }
      else { z += PIO2_2; y[0] = z + PIO2_2T; y[1] = (z - y[0]) + PIO2_2T; }
      return -1;
    }
  }
  if (ix <= 0x413921fb) { /* medium size */
    t = fabs(x);
    n = (int)(t * INVPIO2 + 0.5);
    fn = (double)n;
    r = t - fn * PIO2_1;
    w = fn * PIO2_1T;
    if (n < 32 && ix != npio2_hw[n - 1]) {
      y[0] = r - w;
    } else {
      j = ix >> 20;
      y[0] = r - w;
      i = j - ((hi_word(y[0]) >> 20) & 0x7ff);
      if (i > 16) {
        t = r; w = fn * PIO2_2; r = t - w;
        w = fn * PIO2_2T - ((t - r) - w);
        y[0] = r - w;
        i = j - ((hi_word(y[0]) >> 20) & 0x7ff);
        if (i > 49) {
          t = r; w = fn * PIO2_3; r = t - w;
          w = fn * PIO2_3T - ((t - r) - w);
          y[0] = r - w;
        }
      }
    }
    y[1] = (r - y[0]) - w;
    if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
    return n;
  }
  y[0] = y[1] = NAN;
  return 0x7fffffff;
}

/* ---- kernels on [-pi/4, pi/4] (fdlibm k_sin.c, k_cos.c, k_tan.c) ---- */
#define S1 -1.66666666666666324348e-01 /* 0xBFC55555 55555549 */
#define S2  8.33333333332248946124e-03 /* 0x3F811111 1110F8A6 */
#define S3 -1.98412698298579493134e-04 /* 0xBF2A01A0 19C161D5 */
#define S4  2.75573137070700676789e-06 /* 0x3EC71DE3 57B1FE7D */
#define S5 -2.50507602534068634195e-08 /* 0xBE5AE5E6 8A2B9CEB */
#define S6  1.58969099521155010221e-10 /* 0x3DE5D93A 5ACFD57C */

double or_kernel_sin(double x, double y, int iy) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  double z, r, v;
  if (ix < 0x3e400000) { if ((int)x == 0) return x; }
  z = x * x;
  v = z * x;
  r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

#define C1  4.16666666666666019037e-02 /* 0x3FA55555 5555554C */
#define C2 -1.38888888888741095749e-03 /* 0xBF56C16C 16C15177 */
#define C3  2.48015872894767294178e-05 /* 0x3EFA01A0 19CB1590 */
#define C4 -2.75573143513906633035e-07 /* 0xBE927E4F 809C52AD */
#define C5  2.08757232129817482790e-09 /* 0x3E21EE9E BDB4B1C4 */
#define C6 -1.13596475577881948265e-11 /* 0xBDA8FAE9 BE8838D4 */

double or_kernel_cos(double x, double y) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  double a, hz, z, r, qx;
  if (ix < 0x3e400000) { if ((int)x == 0) return 1.0; }
  z = x * x;
  r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  if (ix > 0x3fe90000) qx = 0.28125;
  else qx = from_words((uint32_t)(ix - 0x00200000), 0);
  hz = 0.5 * z - qx;
  a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

static const double T[13] = {
  3.33333333333334091986e-01, 1.33333333333201242699e-01, 5.39682539762260521377e-02,
  2.18694882948595424599e-02, 8.86323982359930005737e-03, 3.59207910759131235356e-03,
  1.45620945432529025516e-03, 5.88041240820264096874e-04, 2.46463134818469906812e-04,
  7.81794442939557092300e-05, 7.14072491382608190305e-05, -1.85586374855275456654e-05,
  2.59073051863633712884e-05,
};
#define PIO4   7.85398163397448278999e-01 /* 0x3FE921FB 54442D18 */
#define PIO4LO 3.06161699786838301793e-17 /* 0x3C81A626 33145C07 */

double or_kernel_tan(double x, double y, int iy) {
  double z, r, v, w, s;
  int32_t hx = hi_word(x), ix = hx & 0x7fffffff;
  if (ix < 0x3e300000) {
    if ((int)x == 0) {
      if (((ix | (int32_t)lo_word(x)) | (iy + 1)) == 0) return 1.0 / fabs(x);
      if (iy == 1) return x;
      {
        double a, t;
        z = w = x + y;
        z = with_lo(z, 0);
        v = y - (z - x);
        t = a = -1.0 / w;
        t = with_lo(t, 0);
        s = 1.0 + t * z;
        return t + a * (s + t * v);
      }
    }
  }
  if (ix >= 0x3FE59428) {
    if (hx < 0) { x = -x; y = -y; }
    z = PIO4 - x;
    w = PIO4LO - y;
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
  {
    double a, t;
    z = w;
    z = with_lo(z, 0);
    v = r - (z - x);
    t = a = -1.0 / w;
    t = with_lo(t, 0);
    s = 1.0 + t * z;
    return t + a * (s + t * v);
  }
}

double or_sin(double x) {
  double y[2];
  int32_t ix = hi_word(x) & 0x7fffffff;
  int n;
  if (ix <= 0x3fe921fb) return or_kernel_sin(x, 0.0, 0);
  if (ix >= 0x7ff00000) return x - x;
  n = or_rem_pio2(x, y);
  if (n == 0x7fffffff) return sin(x); /* outside the restated range: never on the path */
  switch (n & 3) {
    case 0: return or_kernel_sin(y[0], y[1], 1);
    case 1: return or_kernel_cos(y[0], y[1]);
    case 2: return -or_kernel_sin(y[0], y[1], 1);
    default: return -or_kernel_cos(y[0], y[1]);
  }
}

double or_cos(double x) {
  double y[2];
  int32_t ix = hi_word(x) & 0x7fffffff;
  int n;
  if (ix <= 0x3fe921fb) return or_kernel_cos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  n = or_rem_pio2(x, y);
  if (n == 0x7fffffff) return cos(x);
  switch (n & 3) {
    case 0: return or_kernel_cos(y[0], y[1]);
    case 1: return -or_kernel_sin(y[0], y[1], 1);
    case 2: return -or_kernel_cos(y[0], y[1]);
    default: return or_kernel_sin(y[0], y[1], 1);
  }
}

double or_tan(double x) {
  double y[2];
  int32_t ix = hi_word(x) & 0x7fffffff;
  int n;
  if (ix <= 0x3fe921fb) return or_kernel_tan(x, 0.0, 1);
  if (ix >= 0x7ff00000) return x - x;
  n = or_rem_pio2(x, y);
  if (n == 0x7fffffff) return tan(x);
  return or_kernel_tan(y[0], y[1], 1 - ((n & 1) << 1));
}

/* ---- pow (fdlibm e_pow.c) ---- */
static const double bp[2] = {1.0, 1.5};
static const double dp_h[2] = {0.0, 5.84962487220764160156e-01}; /* 0x3FE2B803 40000000 */
static const double dp_l[2] = {0.0, 1.35003920212974897128e-08}; /* 0x3E4CFDEB 43CFD006 */
#define TWO53  9007199254740992.0
#define HUGE_V 1.0e300
#define TINY_V 1.0e-300
#define L1  5.99999999999994648725e-01 /* 0x3FE33333 33333303 */
#define L2  4.28571428578550184252e-01 /* 0x3FDB6DB6 DB6FABFF */
#define L3  3.33333329818377432918e-01 /* 0x3FD55555 518F264D */
#define L4  2.72728123808534006489e-01 /* 0x3FD17460 A91D4101 */
#define L5  2.30660745775561754067e-01 /* 0x3FCD864A 93C9DB65 */
#define L6  2.06975017800338417784e-01 /* 0x3FCA7E28 4A454EEF */
#define P1  1.66666666666666019037e-01 /* 0x3FC55555 5555553E */
#define P2 -2.77777777770155933842e-03 /* 0xBF66C16C 16BEBD93 */
#define P3  6.61375632143793436117e-05 /* 0x3F11566A AF25DE2C */
#define P4 -1.65339022054652515390e-06 /* 0xBEBBBD41 C5D26BF1 */
#define P5  4.13813679705723846039e-08 /* 0x3E663769 72BEA4D0 */
#define LG2    6.93147180559945286227e-01 /* 0x3FE62E42 FEFA39EF */
#define LG2_H  6.93147182464599609375e-01 /* 0x3FE62E43 00000000 */
#define LG2_L -1.90465429995776804525e-09 /* 0xBE205C61 0CA86C39 */
#define OVT    8.0085662595372944372e-17
#define CP     9.61796693925975554329e-01 /* 0x3FEEC709 DC3A03FD */
#define CP_H   9.61796700954437255859e-01 /* 0x3FEEC709 E0000000 */
#define CP_L  -7.02846165095275826516e-09 /* 0xBE3E2FE0 145B01F5 */
#define IVLN2   1.44269504088896338700e+00 /* 0x3FF71547 652B82FE */
#define IVLN2_H 1.44269502162933349609e+00 /* 0x3FF71547 60000000 */
#define IVLN2_L 1.92596299112661746887e-08 /* 0x3E54AE0B F85DDF44 */

double or_pow(double x, double y) {
  double z, ax, z_h, z_l, p_h, p_l;
  double y1, t1, t2, r, s, t, u, v, w;
  int32_t i, j, k, yisint, n;
  int32_t hx, hy, ix, iy;
  uint32_t lx, ly;

  hx = hi_word(x); lx = lo_word(x);
  hy = hi_word(y); ly = lo_word(y);
  ix = hx & 0x7fffffff; iy = hy & 0x7fffffff;

  if ((iy | (int32_t)ly) == 0) return 1.0;
  if (ix > 0x7ff00000 || ((ix == 0x7ff00000) && (lx != 0)) ||
      iy > 0x7ff00000 || ((iy == 0x7ff00000) && (ly != 0)))
    return x + y;

  yisint = 0;
  if (hx < 0) {
    if (iy >= 0x43400000) yisint = 2;
    else if (iy >= 0x3ff00000) {
      k = (iy >> 20) - 0x3ff;
      if (k > 20) {
        j = (int32_t)(ly >> (52 - k));
        if (((uint32_t)j << (52 - k)) == ly) yisint = 2 - (j & 1);
      } else if (ly == 0) {
        j = iy >> (20 - k);
        if ((j << (20 - k)) == iy) yisint = 2 - (j & 1);
      }
    }
  }

  if (ly == 0) {
    if (iy == 0x7ff00000) {
      if (((ix - 0x3ff00000) | (int32_t)lx) == 0) return y - y;
      else if (ix >= 0x3ff00000) return (hy >= 0) ? y : 0.0;
      else return (hy < 0) ? -y : 0.0;
    }
    if (iy == 0x3ff00000) { if (hy < 0) return 1.0 / x; else return x; }
    if (hy == 0x40000000) return x * x;
    if (hy == 0x3fe00000) { if (hx >= 0) return sqrt(x); }
  }

  ax = fabs(x);
  if (lx == 0) {
    if (ix == 0x7ff00000 || ix == 0 || ix == 0x3ff00000) {
      z = ax;
      if (hy < 0) z = 1.0 / z;
      if (hx < 0) {
        if (((ix - 0x3ff00000) | yisint) == 0) z = (z - z) / (z - z);
        else if (yisint == 1) z = -z;
      }
      return z;
    }
  }

  n = (hx >> 31) + 1;
  if ((n | yisint) == 0) return (x - x) / (x - x);

  s = 1.0;
  if ((n | (yisint - 1)) == 0) s = -1.0;

  if (iy > 0x41e00000) {
    if (iy > 0x43f00000) {
      if (ix <= 0x3fefffff) return (hy < 0) ? HUGE_V * HUGE_V : TINY_V * TINY_V;
      if (ix >= 0x3ff00000) return (hy > 0) ? HUGE_V * HUGE_V : TINY_V * TINY_V;
    }
    if (ix < 0x3fefffff) return (hy < 0) ? s * HUGE_V * HUGE_V : s * TINY_V * TINY_V;
    if (ix > 0x3ff00000) return (hy > 0) ? s * HUGE_V * HUGE_V : s * TINY_V * TINY_V;
    t = ax - 1.0;
    w = (t * t) * (0.5 - t * (0.3333333333333333333333 - t * 0.25));
    u = IVLN2_H * t;
    v = t * IVLN2_L - w * IVLN2;
    t1 = u + v;
    t1 = with_lo(t1, 0);
    t2 = v - (t1 - u);
  } else {
    double ss, s2, s_h, s_l, t_h, t_l;
    n = 0;
    if (ix < 0x00100000) { ax *= TWO53; n -= 53; ix = hi_word(ax); }
    n += ((ix) >> 20) - 0x3ff;
    j = ix & 0x000fffff;
    ix = j | 0x3ff00000;
    if (j <= 0x3988E) k = 0;
    else if (j < 0xBB67A) k = 1;
    else { k = 0; n += 1; ix -= 0x00100000; }
    ax = with_hi(ax, (uint32_t)ix);

    u = ax - bp[k];
    v = 1.0 / (ax + bp[k]);
    ss = u * v;
    s_h = ss;
    s_h = with_lo(s_h, 0);
    t_h = from_words((uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)), 0);
    t_l = ax - (t_h - bp[k]);
    s_l = v * ((u - s_h * t_h) - s_h * t_l);
    s2 = ss * ss;
    r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
    r += s_l * (s_h + ss);
    s2 = s_h * s_h;
    t_h = 3.0 + s2 + r;
    t_h = with_lo(t_h, 0);
    t_l = r - ((t_h - 3.0) - s2);
    u = s_h * t_h;
    v = s_l * t_h + t_l * ss;
    p_h = u + v;
    p_h = with_lo(p_h, 0);
    p_l = v - (p_h - u);
    z_h = CP_H * p_h;
    z_l = CP_L * p_h + p_l * CP + dp_l[k];
    t = (double)n;
    t1 = (((z_h + z_l) + dp_h[k]) + t);
    t1 = with_lo(t1, 0);
    t2 = z_l - (((t1 - t) - dp_h[k]) - z_h);
  }

  y1 = y;
  y1 = with_lo(y1, 0);
  p_l = (y - y1) * t1 + y * t2;
  p_h = y1 * t1;
  z = p_l + p_h;
  j = hi_word(z);
  i = (int32_t)lo_word(z);
  if (j >= 0x40900000) {
    if (((j - 0x40900000) | i) != 0) return s * HUGE_V * HUGE_V;
    if (p_l + OVT > z - p_h) return s * HUGE_V * HUGE_V;
  } else if ((j & 0x7fffffff) >= 0x4090cc00) {
    if (((j - (int32_t)0xc090cc00) | i) != 0) return s * TINY_V * TINY_V;
    if (p_l <= z - p_h) return s * TINY_V * TINY_V;
  }
  i = j & 0x7fffffff;
  k = (i >> 20) - 0x3ff;
  n = 0;
  if (i > 0x3fe00000) {
    n = j + (0x00100000 >> (k + 1));
    k = ((n & 0x7fffffff) >> 20) - 0x3ff;
    t = from_words((uint32_t)(n & ~(0x000fffff >> k)), 0);
    n = ((n & 0x000fffff) | 0x00100000) >> (20 - k);
    if (j < 0) n = -n;
    p_h -= t;
  }
  t = p_l + p_h;
  t = with_lo(t, 0);
  u = t * LG2_H;
  v = (p_l - (t - p_h)) * LG2 + t * LG2_L;
  z = u + v;
  w = v - (z - u);
  t = z * z;
  t1 = z - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  r = (z * t1) / (t1 - 2.0) - (w + z * w);
  z = 1.0 - (r - z);
  j = hi_word(z);
  j += (n << 20);
  if ((j >> 20) <= 0) z = scalbn(z, n);
  else z = with_hi(z, (uint32_t)(hi_word(z) + (n << 20)));
  return s * z;
}

/* Math.hypot(a, b) as V8's builtin computes it */
double or_hypot(double a, double b) {
  double fa = fabs(a), fb = fabs(b), mx, sum = 0, comp = 0, nrm, summand, prelim;
  if (isinf(a) || isinf(b)) return INFINITY;
  if (isnan(a) || isnan(b)) return NAN;
  mx = fa > fb ? fa : fb;
  if (mx == 0) return 0;
  nrm = fa / mx; summand = nrm * nrm - comp; prelim = sum + summand;
  comp = (prelim - sum) - summand; sum = prelim;
  nrm = fb / mx; summand = nrm * nrm - comp; prelim = sum + summand;
  comp = (prelim - sum) - summand; sum = prelim;
  return sqrt(sum) * mx;
}

/* Math.round: nearest integer, halves toward +inf */
double or_js_round(double x) {
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

/* ---- atan (fdlibm s_atan.c), for the lobe diagram's phase (src/sld.js:55) ---- */
static const double atanhi[4] = {
  4.63647609000806093515e-01, 7.85398163397448278999e-01,
  9.82793723247329054082e-01, 1.57079632679489655800e+00,
};
static const double atanlo[4] = {
  2.26987774529616870924e-17, 3.06161699786838301793e-17,
  1.39033110312309984516e-17, 6.12323399573676603587e-17,
};
static const double aT[11] = {
  3.33333333333329318027e-01, -1.99999999998764832476e-01, 1.42857142725034663711e-01,
  -1.11111104054623557880e-01, 9.09088713343650656196e-02, -7.69187620504482999495e-02,
  6.66107313738753120669e-02, -5.83357013379057348645e-02, 4.97687799461593236017e-02,
  -3.65315727442169155270e-02, 1.62858201153657823623e-02,
};

double or_atan(double x) {
  double w, s1, s2, z;
  int32_t ix, hx, id;
  hx = hi_word(x);
  ix = hx & 0x7fffffff;
  if (ix >= 0x44100000) {
    if (ix > 0x7ff00000 || (ix == 0x7ff00000 && (lo_word(x) != 0))) return x + x;
    if (hx > 0) return atanhi[3] + atanlo[3];
    return -atanhi[3] - atanlo[3];
  }
  if (ix < 0x3fdc0000) {
    if (ix < 0x3e200000) return x;
    id = -1;
  } else {
    x = fabs(x);
    if (ix < 0x3ff30000) {
      if (ix < 0x3fe60000) { id = 0; x = (2.0 * x - 1.0) / (2.0 + x); }
      else { id = 1; x = (x - 1.0) / (x + 1.0); }
    } else {
      if (ix < 0x40038000) { id = 2; x = (x - 1.5) / (1.0 + 1.5 * x); }
      else { id = 3; x = -1.0 / x; }
    }
  }
  z = x * x;
  w = z * z;
  s1 = z * (aT[0] + w * (aT[2] + w * (aT[4] + w * (aT[6] + w * (aT[8] + w * aT[10])))));
  s2 = w * (aT[1] + w * (aT[3] + w * (aT[5] + w * (aT[7] + w * aT[9]))));
  if (id < 0) return x - x * (s1 + s2);
  z = atanhi[id] - ((x * (s1 + s2) - atanlo[id]) - x);
  return (hx < 0) ? -z : z;
}

/* Math.atan2 (fdlibm e_atan2.c as FreeBSD's msun keeps it): NaN, x = 1, zeros and infinities
   first, then atan(|y / x|) put in the quadrant with pi's low part; a ratio beyond 2^60 gives
   +-pi/2 whatever the sign of x, and a negative x with a ratio under 2^-60 gives +-pi */
double or_atan2(double y, double x) {
  static const double tiny = 1.0e-300, pi_o_4 = 7.8539816339744827900e-01, pi_o_2 = 1.5707963267948965580e+00,
                      pi = 3.1415926535897931160e+00, pi_lo = 1.2246467991473531772e-16;
  const int32_t hx = hi_word(x), hy = hi_word(y), ix = hx & 0x7fffffff, iy = hy & 0x7fffffff;
  int32_t k, m;
  double z;
  if (isnan(x) || isnan(y)) return x + y;
  if (x == 1.0) return or_atan(y);
  m = ((hy >> 31) & 1) | ((hx >> 30) & 2);   /* 2 sign(x) + sign(y) */
  if (y == 0) {
    if (m < 2) return y;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if (x == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (isinf(x)) {
    if (isinf(y)) {
      if (m == 0) return pi_o_4 + tiny;
      if (m == 1) return -pi_o_4 - tiny;
      if (m == 2) return 3.0 * pi_o_4 + tiny;
      return -3.0 * pi_o_4 - tiny;
    }
    if (m == 0) return 0.0;
    if (m == 1) return -0.0;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if (isinf(y)) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  k = (iy - ix) >> 20;
  if (k > 60) { z = pi_o_2 + 0.5 * pi_lo; m &= 1; }
  else if (hx < 0 && k < -60) z = 0.0;
  else z = or_atan(fabs(y / x));
  if (m == 0) return z;
  if (m == 1) return -z;
  if (m == 2) return pi - (z - pi_lo);
  return (z - pi_lo) - pi;
}
