/*
 * oracle/sld.c — TEST INFRASTRUCTURE ONLY (the oracle).  Restates the zero-order
 * stability-lobe solution of src/sld.js (Altintas & Budak) and the FRF of src/frf.js:
 *
 *   frf(modes, f)          src/frf.js:14-21
 *   directionalCoeffs      src/sld.js:6-15
 *   stabilityLobes         src/sld.js:23-61
 *   alimAt                 src/sld.js:64-76
 *
 * Same operation order as the JavaScript; Math.* from fdlibm.c.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define JS_PI 3.141592653589793

static void frf(const mc_mode *m, int n, double f, double *re, double *im) {
  int i;
  *re = 0; *im = 0;
  for (i = 0; i < n; i++) {
    double r = f / m[i].fn, a = 1 - r * r, b = 2 * m[i].zeta * r, d = m[i].k * (a * a + b * b);
    *re += a / d;
    *im -= b / d;
  }
}

void or_directional_coeffs(double Kr, double phiSt, double phiEx, double out[4]) {
  double p, a[4], b[4];
  int t;
  for (t = 0; t < 2; t++) {
    double *o = t ? b : a;
    p = t ? phiSt : phiEx;
    o[0] = 0.5 * (or_cos(2 * p) - 2 * Kr * p + Kr * or_sin(2 * p));
    o[1] = 0.5 * (-or_sin(2 * p) - 2 * p + Kr * or_cos(2 * p));
    o[2] = 0.5 * (-or_sin(2 * p) + 2 * p + Kr * or_cos(2 * p));
    o[3] = 0.5 * (-or_cos(2 * p) - 2 * Kr * p - Kr * or_sin(2 * p));
  }
  for (t = 0; t < 4; t++) out[t] = a[t] - b[t];
}

static double js_sign(double v) {
  if (isnan(v)) return NAN;
  if (v > 0) return 1;
  if (v < 0) return -1;
  return v;
}

typedef struct { double wc, alim, psi; } Pt;
typedef struct { double rpm, alim; } LP;

static int cmp_lp(const void *a, const void *b) {
  double x = ((const LP *)a)->rpm, y = ((const LP *)b)->rpm;
  return x < y ? -1 : x > y ? 1 : 0;
}

int or_stability_lobes(int N, double Ktc, double Krc, double phiSt, double phiEx, int lobes,
                       double rpmMin, double rpmMax, const mc_mode *mx, int nmx,
                       const mc_mode *my, int nmy, double *rpm, double *alim, int max_pts,
                       int *lobe_len) {
  double al[4], det, f0, f1, fmin = INFINITY, fmax = -INFINITY;
  int NS, s, j, i, total = 0, npts = 0, nout = 0;
  Pt *pts;
  LP *tmp;
  or_directional_coeffs(Krc / Ktc, phiSt, phiEx, al);
  det = al[0] * al[3] - al[1] * al[2];
  for (i = 0; i < nmx; i++) { if (mx[i].fn < fmin) fmin = mx[i].fn; if (mx[i].fn > fmax) fmax = mx[i].fn; }
  for (i = 0; i < nmy; i++) { if (my[i].fn < fmin) fmin = my[i].fn; if (my[i].fn > fmax) fmax = my[i].fn; }
  f0 = 0.8 * fmin; f1 = 1.7 * fmax;
  NS = 400 * (nmx + nmy - 1 < 4 ? nmx + nmy - 1 : 4);
  pts = (Pt *)malloc(sizeof(Pt) * (NS > 0 ? NS : 1));
  tmp = (LP *)malloc(sizeof(LP) * (NS > 0 ? NS : 1));
  for (s = 0; s < NS; s++) {
    double fc = f0 + (f1 - f0) * s / (NS - 1), wc = 2 * JS_PI * fc;
    double Px[2], Py[2], pp[2], a0[2], a1[2], sq[2], disc[2], bestA = 0, bestK = 0;
    int has = 0, g;
    frf(mx, nmx, fc, &Px[0], &Px[1]);
    frf(my, nmy, fc, &Py[0], &Py[1]);
    pp[0] = Px[0] * Py[0] - Px[1] * Py[1];
    pp[1] = Px[0] * Py[1] + Px[1] * Py[0];
    a0[0] = pp[0] * det - pp[1] * 0;
    a0[1] = pp[0] * 0 + pp[1] * det;
    a1[0] = al[0] * Px[0] + al[3] * Py[0];
    a1[1] = al[0] * Px[1] + al[3] * Py[1];
    sq[0] = (a1[0] * a1[0] - a1[1] * a1[1]) - 4 * a0[0];
    sq[1] = (a1[0] * a1[1] + a1[1] * a1[0]) - 4 * a0[1];
    {
      double r = or_hypot(sq[0], sq[1]);
      double sgn = js_sign((sq[1] != 0 && !isnan(sq[1])) ? sq[1] : 1);
      double t = (r - sq[0]) / 2;
      disc[0] = sqrt((r + sq[0]) / 2);
      disc[1] = sgn * sqrt(t > 0 ? t : 0);
    }
    for (g = 0; g < 2; g++) {
      double sg = g ? -1 : 1, num[2], den[2], d, L[2], kap, a;
      num[0] = -(a1[0] + sg * disc[0]);
      num[1] = -(a1[1] + sg * disc[1]);
      den[0] = 2 * a0[0];
      den[1] = 2 * a0[1];
      d = den[0] * den[0] + den[1] * den[1];
      L[0] = (num[0] * den[0] + num[1] * den[1]) / d;
      L[1] = (num[1] * den[0] - num[0] * den[1]) / d;
      if (L[0] >= 0) continue;
      kap = L[1] / L[0];
      a = -2 * JS_PI * L[0] * (1 + kap * kap) / (N * Ktc);
      if (a > 0 && (!has || a < bestA)) { bestA = a; bestK = kap; has = 1; }
    }
    if (has) { pts[npts].wc = wc; pts[npts].alim = bestA; pts[npts].psi = or_atan(bestK); npts++; }
  }
  for (j = 0; j < lobes; j++) {
    int n = 0;
    for (i = 0; i < npts; i++) {
      double eps = JS_PI - 2 * pts[i].psi;
      double T = (eps + 2 * j * JS_PI) / pts[i].wc;
      double r = 60 / (N * T);
      if (r >= rpmMin * 0.8 && r <= rpmMax * 1.2) { tmp[n].rpm = r; tmp[n].alim = pts[i].alim; n++; }
    }
    /* Array.prototype.sort is stable; ties in rpm keep insertion order */
    for (i = 1; i < n; i++) {
      LP v = tmp[i];
      int k = i - 1;
      while (k >= 0 && tmp[k].rpm > v.rpm) { tmp[k + 1] = tmp[k]; k--; }
      tmp[k + 1] = v;
    }
    if (n) {
      for (i = 0; i < n && total < max_pts; i++) { rpm[total] = tmp[i].rpm; alim[total] = tmp[i].alim; total++; }
      lobe_len[nout++] = n;
    }
  }
  for (j = nout; j < lobes; j++) lobe_len[j] = 0;
  (void)cmp_lp;
  free(pts);
  free(tmp);
  return total;
}

double or_alim_at(const double *rpm, const double *alim, const int *lobe_len, int nlobes, double S) {
  double best = INFINITY;
  int j, i, off = 0;
  for (j = 0; j < nlobes; j++) {
    for (i = 1; i < lobe_len[j]; i++) {
      double ar = rpm[off + i - 1], br = rpm[off + i];
      if (S >= ar && S <= br) {
        double d = br - ar;
        double t = (S - ar) / (d != 0 ? d : 1);
        double v = alim[off + i - 1] + (alim[off + i] - alim[off + i - 1]) * t;
        if (v < best) best = v;
      }
    }
    off += lobe_len[j];
  }
  return best;
}

/* The composite curve of Simulator.lobes for a speed-dependent tool tip (src/sim.js:270-281):
 * node i's diagram (stabilityLobes over [S0, S0 + node] with that node's modes) read by alimAt
 * at S = S0, S0 + step, ... while S < min(hi, S0 + node); a speed no lobe covers reads 1e3.
 * Returns the points written (all of them, however large, when rpm is NULL). */
int or_composite_lobes(int n, const int *N, const double *Ktc, const double *Krc, const double *phiSt,
                       const double *phiEx, const int *lobes, const double *S0, const double *S1,
                       const mc_mode *mx, const int *nmx, const mc_mode *my, const int *nmy, int mstride,
                       double step, double hi, double *rpm, double *alim) {
  const int cap = 1 << 20;
  double *lr = malloc(sizeof(double) * cap), *la = malloc(sizeof(double) * cap);
  int *ll = malloc(sizeof(int) * 100000);
  int i, total = 0;
  for (i = 0; i < n; i++) {
    const double top = hi < S1[i] ? hi : S1[i];
    double S;
    int nl = 0, j;
    if (rpm) {
      or_stability_lobes(N[i], Ktc[i], Krc[i], phiSt[i], phiEx[i], lobes[i], S0[i], S1[i], mx + i * mstride, nmx[i],
                         my + i * mstride, nmy[i], lr, la, cap, ll);
      for (j = 0; j < lobes[i]; j++) if (ll[j]) nl = j + 1;
    }
    for (S = S0[i]; S < top; S += step) {
      if (rpm) {
        const double a = or_alim_at(lr, la, ll, nl, S);
        rpm[total] = S;
        alim[total] = a == INFINITY ? 1e3 : a;
      }
      total++;
    }
  }
  free(lr);
  free(la);
  free(ll);
  return total;
}
