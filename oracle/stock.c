/*
 * oracle/stock.c — TEST INFRASTRUCTURE ONLY (the oracle).  Restates the Z-heightfield
 * stock of src/stock.js:
 *
 *   constructor     src/stock.js:9-40   (H = zTop = 0, zBottom = -h, wallD = +inf)
 *   stamp           src/stock.js:118-164 (flat / ball; no LUT, no feed marks)
 *   sample          src/stock.js:85-92
 *
 * FP64 arithmetic with Float32Array / Int32Array stores, same loop order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define EDGE_TIE (2 * 0.0015)

struct or_stock {
  double w, l, zTop, zBottom, hx, hy, wallBand, removed;
  int32_t nx, ny;
  float *H, *err, *wallD;
  int32_t *lineCut, *lineWall;
};

or_stock *or_stock_new(double w, double l, double h, double cell) {
  or_stock *s = (or_stock *)calloc(1, sizeof(or_stock));
  size_t n, k;
  if (!s) return NULL;
  s->w = w; s->l = l; s->zTop = 0; s->zBottom = -h;
  s->nx = (int32_t)or_js_round(w / cell) + 1;
  s->ny = (int32_t)or_js_round(l / cell) + 1;
  s->hx = w / (s->nx - 1);
  s->hy = l / (s->ny - 1);
  n = (size_t)s->nx * s->ny;
  s->H = (float *)calloc(n, sizeof(float));
  s->err = (float *)calloc(n, sizeof(float));
  s->wallD = (float *)malloc(n * sizeof(float));
  s->lineCut = (int32_t *)calloc(n, sizeof(int32_t));
  s->lineWall = (int32_t *)calloc(n, sizeof(int32_t));
  for (k = 0; k < n; k++) { s->H[k] = (float)s->zTop; s->wallD[k] = INFINITY; }
  return s;
}

void or_stock_free(or_stock *s) {
  if (!s) return;
  free(s->H); free(s->err); free(s->wallD); free(s->lineCut); free(s->lineWall); free(s);
}

void or_stock_dims(const or_stock *s, int32_t *nx, int32_t *ny, double *hx, double *hy) {
  *nx = s->nx; *ny = s->ny; *hx = s->hx; *hy = s->hy;
}
void or_stock_set_wall_band(or_stock *s, double band) { s->wallBand = band; }
float *or_stock_h(or_stock *s) { return s->H; }
float *or_stock_err(or_stock *s) { return s->err; }
float *or_stock_walld(or_stock *s) { return s->wallD; }
int32_t *or_stock_linecut(or_stock *s) { return s->lineCut; }
int32_t *or_stock_linewall(or_stock *s) { return s->lineWall; }

static double dmax(double a, double b) { return a > b ? a : b; }

double or_stock_stamp(or_stock *s, double cx, double cy, double tipZ, double R, int type,
                      double errVal, int line) {
  double Rw, Rw2, R2, zRim, vol = 0;
  int32_t i0, i1, j0, j1, i, j, n = s->nx;
  if (tipZ >= s->zTop) return 0;
  Rw = R + dmax(2 * dmax(s->hx, s->hy), s->wallBand);
  Rw2 = Rw * Rw;
  {
    double a = floor((cx - Rw) / s->hx), b = ceil((cx + Rw) / s->hx);
    double c = floor((cy - Rw) / s->hy), d = ceil((cy + Rw) / s->hy);
    i0 = (int32_t)dmax(0, a); i1 = (int32_t)(b < s->nx - 1 ? b : s->nx - 1);
    j0 = (int32_t)dmax(0, c); j1 = (int32_t)(d < s->ny - 1 ? d : s->ny - 1);
    if (a > 2e9 || c > 2e9) return 0;
  }
  if (i0 > i1 || j0 > j1) return 0;
  R2 = R * R;
  zRim = tipZ + (type == MC_TOOL_BALL ? R : 0);
  for (j = j0; j <= j1; j++) {
    double dy = j * s->hy - cy;
    for (i = i0; i <= i1; i++) {
      double dx = i * s->hx - cx;
      double r2 = dx * dx + dy * dy, z;
      int32_t k = j * n + i;
      if (r2 >= R2 - EDGE_TIE * R) {
        if (r2 <= Rw2 && s->H[k] > zRim) {
          double dd = sqrt(r2) - R;
          if (dd < s->wallD[k]) { s->wallD[k] = (float)dd; s->lineWall[k] = line; }
        }
        continue;
      }
      z = type == MC_TOOL_BALL ? tipZ + R - sqrt(R2 - r2) : tipZ;
      if (s->H[k] > z) {
        vol += s->H[k] - z;
        s->H[k] = (float)dmax(z, s->zBottom);
        s->err[k] = (float)errVal;
        s->wallD[k] = INFINITY;
        s->lineCut[k] = line;
      }
    }
  }
  vol *= s->hx * s->hy;
  s->removed += vol;
  return vol;
}

void or_stamp_batch(or_stock *s, const mc_stamp *st, int n, double *vol) {
  int i;
  for (i = 0; i < n; i++) {
    double v = or_stock_stamp(s, st[i].cx, st[i].cy, st[i].tipZ, st[i].R, st[i].type,
                              st[i].errVal, st[i].line);
    if (vol) vol[i] = v;
  }
}

double or_stock_sample(const or_stock *s, double x, double y) {
  double fx = x / s->hx, fy = y / s->hy, tx, ty, a, b, c, d;
  int32_t i, j, n = s->nx;
  if (fx < 0 || fy < 0 || fx > s->nx - 1 || fy > s->ny - 1) return -INFINITY;
  i = (int32_t)fx; if (i > s->nx - 2) i = s->nx - 2;
  j = (int32_t)fy; if (j > s->ny - 2) j = s->ny - 2;
  tx = fx - i; ty = fy - j;
  a = s->H[j * n + i]; b = s->H[j * n + i + 1]; c = s->H[(j + 1) * n + i]; d = s->H[(j + 1) * n + i + 1];
  return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}
