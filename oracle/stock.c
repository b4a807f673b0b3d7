/*
 * oracle/stock.c — TEST INFRASTRUCTURE ONLY (the oracle).  Restates the Z-heightfield
 * stock of src/stock.js and the tip table of src/assembly.js:
 *
 *   constructor     src/stock.js:9-40    (H = zTop = 0, zBottom = -h, wallD = +inf,
 *                                         8x8 block maxima, all blocks dirty)
 *   markBlocks      src/stock.js:41-44
 *   blockMax        src/stock.js:45-54   (lazy, recomputed when dirty)
 *   maxInDisk       src/stock.js:57-80   (serial block walk)
 *   clearDirty      src/stock.js:81
 *   sample          src/stock.js:85-92
 *   markWall        src/stock.js:101-116
 *   stamp           src/stock.js:118-164 (flat / ball / profile table, feed marks,
 *                                         dirty rectangle and block marks)
 *   tipLut          src/assembly.js:38-54
 *
 * FP64 arithmetic with Float32Array / Int32Array stores, same loop order.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define EDGE_TIE (2 * 0.0015)
#define LUT_DR 0.02
#define BLK 8

struct or_stock {
  double w, l, zTop, zBottom, hx, hy, wallBand, removed;
  int32_t nx, ny, bnx, bny;
  float *H, *err, *wallD, *markS, *markF, *bmax, *wallErr;
  int32_t *lineCut, *lineWall;
  unsigned char *bdirty;
  double di0, dj0, di1, dj1;   /* dirty rectangle: Infinity / -1 when empty */
};

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

void or_stock_clear_dirty(or_stock *s) {
  s->di0 = INFINITY; s->dj0 = INFINITY; s->di1 = -1; s->dj1 = -1;
}

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
  s->markS = (float *)calloc(n, sizeof(float));
  s->markF = (float *)calloc(n, sizeof(float));
  s->wallErr = (float *)calloc(n, sizeof(float));
  s->lineCut = (int32_t *)calloc(n, sizeof(int32_t));
  s->lineWall = (int32_t *)calloc(n, sizeof(int32_t));
  for (k = 0; k < n; k++) { s->H[k] = (float)s->zTop; s->wallD[k] = INFINITY; }
  s->bnx = (s->nx + BLK - 1) / BLK;
  s->bny = (s->ny + BLK - 1) / BLK;
  s->bmax = (float *)calloc((size_t)s->bnx * s->bny, sizeof(float));
  s->bdirty = (unsigned char *)malloc((size_t)s->bnx * s->bny);
  memset(s->bdirty, 1, (size_t)s->bnx * s->bny);
  or_stock_clear_dirty(s);
  return s;
}

void or_stock_free(or_stock *s) {
  if (!s) return;
  free(s->H); free(s->err); free(s->wallD); free(s->markS); free(s->markF); free(s->wallErr);
  free(s->lineCut); free(s->lineWall); free(s->bmax); free(s->bdirty); free(s);
}

void or_stock_dims(const or_stock *s, int32_t *nx, int32_t *ny, double *hx, double *hy) {
  *nx = s->nx; *ny = s->ny; *hx = s->hx; *hy = s->hy;
}
void or_stock_set_wall_band(or_stock *s, double band) { s->wallBand = band; }
double or_stock_zbottom(const or_stock *s) { return s->zBottom; }
float *or_stock_h(or_stock *s) { return s->H; }
float *or_stock_err(or_stock *s) { return s->err; }
float *or_stock_walld(or_stock *s) { return s->wallD; }
float *or_stock_marks(or_stock *s) { return s->markS; }
float *or_stock_markf(or_stock *s) { return s->markF; }
float *or_stock_wallerr(or_stock *s) { return s->wallErr; }

/* markWall (src/stock.js:101-116): val into wallErr on the cells within two cells outside
 * the cutter on the side facing (nx, ny); the box joins the dirty rectangle */
void or_stock_mark_wall(or_stock *s, double cx, double cy, double R, double nx, double ny,
                        double val) {
  const double w = 2 * dmax(s->hx, s->hy), Ro = R + w;
  double a = floor((cx - Ro) / s->hx), b = ceil((cx + Ro) / s->hx);
  double c = floor((cy - Ro) / s->hy), d = ceil((cy + Ro) / s->hy);
  int32_t i, j;
  a = dmax(0, a); b = dmin(s->nx - 1, b); c = dmax(0, c); d = dmin(s->ny - 1, d);
  for (j = (int32_t)c; j <= (int32_t)d && c <= d; j++)
    for (i = (int32_t)a; i <= (int32_t)b && a <= b; i++) {
      const double dx = i * s->hx - cx, dy = j * s->hy - cy, r = or_hypot(dx, dy);
      if (r <= R || r > Ro || (dx * nx + dy * ny) < 0.5 * r) continue;
      s->wallErr[j * s->nx + i] = (float)val;
    }
  if (a <= b && c <= d) {
    s->di0 = dmin(s->di0, a); s->dj0 = dmin(s->dj0, c);
    s->di1 = dmax(s->di1, b); s->dj1 = dmax(s->dj1, d);
  }
}
int32_t *or_stock_linecut(or_stock *s) { return s->lineCut; }
int32_t *or_stock_linewall(or_stock *s) { return s->lineWall; }

/* rect[4] = {i0, j0, i1, j1}; empty: {INT32_MAX, INT32_MAX, -1, -1} */
void or_stock_dirty(const or_stock *s, int32_t *rect) {
  rect[0] = isinf(s->di0) ? INT32_MAX : (int32_t)s->di0;
  rect[1] = isinf(s->dj0) ? INT32_MAX : (int32_t)s->dj0;
  rect[2] = (int32_t)s->di1;
  rect[3] = (int32_t)s->dj1;
}

/* src/stock.js:41-44 */
static void mark_blocks(or_stock *s, int32_t i0, int32_t j0, int32_t i1, int32_t j1) {
  int32_t bi, bj;
  for (bj = j0 / BLK; bj <= j1 / BLK; bj++)
    for (bi = i0 / BLK; bi <= i1 / BLK; bi++) s->bdirty[bj * s->bnx + bi] = 1;
}

/* src/stock.js:45-54 */
float or_stock_block_max(or_stock *s, int32_t bi, int32_t bj) {
  int32_t k = bj * s->bnx + bi, i, j, n = s->nx;
  if (s->bdirty[k]) {
    float m = -INFINITY;
    for (j = bj * BLK; j < (s->ny < bj * BLK + BLK ? s->ny : bj * BLK + BLK); j++)
      for (i = bi * BLK; i < (n < bi * BLK + BLK ? n : bi * BLK + BLK); i++)
        if (s->H[j * n + i] > m) m = s->H[j * n + i];
    s->bmax[k] = m;
    s->bdirty[k] = 0;
  }
  return s->bmax[k];
}

/* src/stock.js:57-80: out = {h, x, y} */
void or_stock_max_in_disk(or_stock *s, double cx, double cy, double r, double above, double *out) {
  const double hx = s->hx, hy = s->hy;
  const int32_t n = s->nx;
  double a = floor((cx - r) / hx), b = ceil((cx + r) / hx), c = floor((cy - r) / hy),
         d = ceil((cy + r) / hy);
  double best = -INFINITY, bx = 0, by = 0, r2;
  int32_t i0, i1, j0, j1, bi, bj, i, j;
  a = dmax(0, a); b = dmin(s->nx - 1, b); c = dmax(0, c); d = dmin(s->ny - 1, d);
  out[0] = best; out[1] = bx; out[2] = by;
  if (a > b || c > d) return;
  i0 = (int32_t)a; i1 = (int32_t)b; j0 = (int32_t)c; j1 = (int32_t)d;
  r2 = r * r;
  for (bj = j0 / BLK; bj <= j1 / BLK; bj++)
    for (bi = i0 / BLK; bi <= i1 / BLK; bi++) {
      const double m = or_stock_block_max(s, bi, bj);
      double xa, xb, ya, yb, far;
      int32_t ja, jb, ia, ib;
      if (m <= dmax(best, above)) continue;
      xa = (double)(bi * BLK) * hx - cx;
      xb = (double)(n - 1 < bi * BLK + BLK - 1 ? n - 1 : bi * BLK + BLK - 1) * hx - cx;
      ya = (double)(bj * BLK) * hy - cy;
      yb = (double)(s->ny - 1 < bj * BLK + BLK - 1 ? s->ny - 1 : bj * BLK + BLK - 1) * hy - cy;
      far = dmax(xa * xa, xb * xb) + dmax(ya * ya, yb * yb);
      if (far <= r2) { best = m; bx = cx + (xa + xb) / 2; by = cy + (ya + yb) / 2; continue; }
      ja = j0 > bj * BLK ? j0 : bj * BLK;
      jb = j1 < bj * BLK + BLK - 1 ? j1 : bj * BLK + BLK - 1;
      ia = i0 > bi * BLK ? i0 : bi * BLK;
      ib = i1 < bi * BLK + BLK - 1 ? i1 : bi * BLK + BLK - 1;
      for (j = ja; j <= jb; j++) {
        const double dy = j * hy - cy;
        for (i = ia; i <= ib; i++) {
          const double dx = i * hx - cx, h = s->H[j * n + i];
          if (h > best && dx * dx + dy * dy <= r2) { best = h; bx = i * hx; by = j * hy; }
        }
      }
    }
  out[0] = best; out[1] = bx; out[2] = by;
}

double or_stock_stamp_ex(or_stock *s, double cx, double cy, double tipZ, double R, int type,
                         double errVal, int line, const double *mark, const float *lut,
                         int lutn) {
  double Rw, Rw2, R2, zRim, mf, vol = 0;
  int32_t i0, i1, j0, j1, i, j, n = s->nx, touched = 0;
  if (tipZ >= s->zTop) return 0;
  Rw = R + dmax(2 * dmax(s->hx, s->hy), s->wallBand);
  Rw2 = Rw * Rw;
  {
    double a = floor((cx - Rw) / s->hx), b = ceil((cx + Rw) / s->hx);
    double c = floor((cy - Rw) / s->hy), d = ceil((cy + Rw) / s->hy);
    a = dmax(0, a); b = dmin(s->nx - 1, b); c = dmax(0, c); d = dmin(s->ny - 1, d);
    if (a > b || c > d) return 0;
    i0 = (int32_t)a; i1 = (int32_t)b; j0 = (int32_t)c; j1 = (int32_t)d;
  }
  R2 = R * R;
  mf = mark && mark[2] > 0 && type == MC_TOOL_FLAT && !lut ? mark[2] : 0;
  zRim = tipZ + (lut ? (double)lut[lutn - 1] : type == MC_TOOL_BALL ? R : 0);
  for (j = j0; j <= j1; j++) {
    double dy = j * s->hy - cy;
    for (i = i0; i <= i1; i++) {
      double dx = i * s->hx - cx;
      double r2 = dx * dx + dy * dy, z;
      int32_t k = j * n + i;
      if (r2 >= R2 - EDGE_TIE * R) {
        if (r2 <= Rw2 && s->H[k] > zRim) {
          double dd = sqrt(r2) - R;
          if (dd < s->wallD[k]) { s->wallD[k] = (float)dd; s->lineWall[k] = line; touched = 1; }
        }
        continue;
      }
      if (lut) {
        double q = or_js_round(sqrt(r2) / LUT_DR);
        z = tipZ + (double)lut[(int32_t)(q < lutn - 1 ? q : lutn - 1)];
      } else {
        z = type == MC_TOOL_BALL ? tipZ + R - sqrt(R2 - r2) : tipZ;
      }
      if (s->H[k] > z) {
        vol += s->H[k] - z;
        s->H[k] = (float)dmax(z, s->zBottom);
        s->err[k] = (float)errVal;
        s->wallD[k] = INFINITY;
        s->lineCut[k] = line;
        touched = 1;
      }
      if (mf != 0 && s->H[k] > z - 0.002) {
        const double xs = i * s->hx, ys = j * s->hy, p = -dx * mark[1] + dy * mark[0];
        s->markS[k] = (float)(xs * mark[0] + ys * mark[1] + sqrt(dmax(0, R2 - p * p)));
        s->markF[k] = (float)mf;
        touched = 1;
      }
    }
  }
  if (touched) {
    s->di0 = dmin(s->di0, i0); s->dj0 = dmin(s->dj0, j0);
    s->di1 = dmax(s->di1, i1); s->dj1 = dmax(s->dj1, j1);
    mark_blocks(s, i0, j0, i1, j1);
  }
  vol *= s->hx * s->hy;
  s->removed += vol;
  return vol;
}

double or_stock_stamp(or_stock *s, double cx, double cy, double tipZ, double R, int type,
                      double errVal, int line) {
  return or_stock_stamp_ex(s, cx, cy, tipZ, R, type, errVal, line, NULL, NULL, 0);
}

void or_stamp_batch(or_stock *s, const mc_stamp *st, int n, double *vol) {
  or_stamp_batch_ex(s, st, NULL, n, NULL, NULL, NULL, vol);
}

void or_stamp_batch_ex(or_stock *s, const mc_stamp *st, const mc_stamp_opt *opt, int n,
                       const float *luts, const int32_t *lut_off, const int32_t *lut_len,
                       double *vol) {
  int i;
  for (i = 0; i < n; i++) {
    double mark[3], v;
    const double *mk = NULL;
    const float *lut = NULL;
    int lutn = 0;
    if (opt) {
      if (opt[i].f > 0) { mark[0] = opt[i].ux; mark[1] = opt[i].uy; mark[2] = opt[i].f; mk = mark; }
      if (opt[i].lut >= 0) { lut = luts + lut_off[opt[i].lut]; lutn = lut_len[opt[i].lut]; }
    }
    v = or_stock_stamp_ex(s, st[i].cx, st[i].cy, st[i].tipZ, st[i].R, st[i].type, st[i].errVal,
                          st[i].line, mk, lut, lutn);
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

/* tipLut (src/assembly.js:38-54); returns n, writes out[0..n) when out != NULL */
int or_tip_lut(int shape, double D, double cornerR, const double *er, int en, double edz,
               float *out) {
  const double R = D / 2;
  const int n = (int)ceil(R / LUT_DR) + 1;
  int i, k;
  if (!out) return n;
  for (i = 0; i < n; i++) {
    const double r = dmin(R, i * LUT_DR);
    double z = 0;
    if (shape == MC_TIP_EDGE) {
      z = edz * (en - 1);
      for (k = 0; k < en; k++)
        if (er[k] >= r - 1e-6) {
          z = k == 0 ? 0 : (k - 1 + (r - er[k - 1]) / dmax(1e-9, er[k] - er[k - 1])) * edz;
          break;
        }
    } else if (shape == MC_TIP_BALL) {
      z = R - sqrt(dmax(0, R * R - r * r));
    } else if (shape == MC_TIP_BULL && cornerR > 0) {
      const double rc = dmin(R, cornerR), r0 = R - rc;
      z = r <= r0 ? 0 : rc - sqrt(dmax(0, rc * rc - (r - r0) * (r - r0)));
    }
    out[i] = (float)z;
  }
  return n;
}
