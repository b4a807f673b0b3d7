/*
 * oracle/wall.c — TEST INFRASTRUCTURE ONLY (the oracle).  CPU restatement of the thin-wall
 * model the simulator asks at every cutting sub-step with workpieceFlex on
 * (src/sim.js:407), serially in the reference's order:
 *
 *   wallAt       src/wall.js:11-34   contact point from the engaged bins at the load height,
 *                                    the thinnest crossing within +-40 degrees of the contact
 *                                    normal in 5 degree steps
 *   thickness    src/wall.js:36-42   bilinear samples outward in half-cell steps, up to 20 mm
 *   cellAt       src/wall.js:46-50   the nearest node's height
 *   thicknessAt  src/wall.js:54-60   back out of the material, then across, node by node
 *   measure      src/wall.js:62-112  far floor, free height, load height, length along the
 *                                    wall, the thickness profile, compliance of the
 *                                    cantilever strip, the tip-equivalent mass, the 5 kHz cap
 *
 * Every step length is accumulated by repeated addition, as the reference's loops do, and
 * every expression keeps the reference's operation order.  Math.sin / cos / atan2 / hypot /
 * round are the restated fdlibm (fdlibm.c); x ** 2 is x * x (V8's pow gives exactly that);
 * x ** 3 is the restated fdlibm pow, which differs from Node 12's Math.pow by one ulp on a
 * few per cent of inputs, so k, m, fn and I can differ from the reference's in the last bits
 * (the geometry never does).
 */
#include <math.h>

#include "oracle.h"

#define PI 3.141592653589793
#define NZ 10

static double jmin(double a, double b) { return a < b ? a : b; }   /* Math.min, no NaNs here */
static double jmax(double a, double b) { return a > b ? a : b; }

static double sample(const or_stock *s, double x, double y) { return or_stock_sample(s, x, y); }

/* Stock.floor (src/stock.js:94-97) */
static double floor_at(const or_stock *s, double x, double y) {
  const double v = or_stock_sample(s, x, y);
  return v == -INFINITY ? or_stock_zbottom(s) : v;
}

static double half_step(const or_stock *s) {
  int32_t nx, ny;
  double hx, hy;
  or_stock_dims(s, &nx, &ny, &hx, &hy);
  return jmin(hx, hy) * 0.5;
}

/* src/wall.js:36-42 */
static double thickness(const or_stock *s, double cx, double cy, double nx, double ny, double zLoad) {
  const double step = half_step(s);
  double t;
  for (t = step; t <= 20; t += step)
    if (sample(s, cx + nx * t, cy + ny * t) < zLoad) return t;
  return -1;
}

/* src/wall.js:46-50 */
static double cell_at(const or_stock *s, double x, double y) {
  int32_t nx, ny;
  double hx, hy, i, j;
  or_stock_dims(s, &nx, &ny, &hx, &hy);
  i = or_js_round(x / hx);
  j = or_js_round(y / hy);
  if (i < 0 || j < 0 || i >= nx || j >= ny) return -INFINITY;
  return or_stock_h((or_stock *)s)[(int64_t)j * nx + (int64_t)i];
}

/* src/wall.js:54-60 */
static double thickness_at(const or_stock *s, double cx, double cy, double nx, double ny, double z) {
  const double step = half_step(s);
  double s0 = 0, t, lim;
  while (s0 > -6 && cell_at(s, cx + nx * s0, cy + ny * s0) >= z) s0 -= step;
  lim = s0 + 20;
  for (t = s0 + step; t <= lim; t += step)
    if (cell_at(s, cx + nx * t, cy + ny * t) < z) return t - s0;
  return -1;
}

/* src/wall.js:62-112; returns 1 with *w filled, 0 for null */
static int measure(const or_stock *s, double pz, double cx, double cy, double R, double nx, double ny, double t,
                   double zLoad, double E, double rho, double kTool, mc_wall_out *w) {
  const double step = half_step(s);
  double far = INFINITY, base, mx, my, top, hWall, a, tx, ty, L = 0, bEff, tRef, ts[NZ], comp = 0, k, I, tMean = 0, m,
         fn, q0, lim;
  int q, sg;
  lim = t + 3;
  for (q0 = t; q0 <= lim; q0 += step) far = jmin(far, floor_at(s, cx + nx * q0, cy + ny * q0));
  base = jmax(pz, far);
  mx = cx + nx * t / 2;
  my = cy + ny * t / 2;
  top = sample(s, mx, my);
  hWall = top - base;
  a = jmin(jmax(zLoad - base, 0.3), jmax(hWall, 0.3));
  if (hWall < 0.5 || t < 2 * step * 2) return 0;
  tx = -ny; ty = nx;
  for (sg = 1; sg >= -1; sg -= 2) {
    double d;
    for (d = step; d <= 60; d += step) {
      if (sample(s, mx + tx * d * sg, my + ty * d * sg) < zLoad) break;
      L += step;
    }
  }
  if (L < 2 * t) return 0;
  bEff = jmin(L, 2 * a + t + R);
  tRef = thickness_at(s, cx, cy, nx, ny, zLoad);
  for (q = 0; q < NZ; q++) {
    const double zq = base + (q + 0.5) / NZ * a, tq = thickness_at(s, cx, cy, nx, ny, zq);
    ts[q] = tq > 0 && tRef > 0 ? jmax(t * tq / tRef, t) : t;
  }
  for (q = 0; q < NZ; q++) {
    const double z = (q + 0.5) / NZ * a, Iq = bEff * or_pow(ts[q], 3) / 12;
    comp += (a - z) * (a - z) / (E * Iq) * (a / NZ);
  }
  k = 1 / comp;
  I = bEff * or_pow(t, 3) / 12;
  if (k > 20 * kTool) return 0;
  for (q = 0; q < NZ; q++) tMean = tMean + ts[q];
  tMean = tMean / NZ;
  m = 0.24 * rho * 1e-9 * tMean * bEff * hWall;
  fn = sqrt(k * 1e3 / m) / (2 * PI);
  if (fn > MC_WALL_FN_CAP) {
    const double wc = 2 * PI * MC_WALL_FN_CAP;
    fn = MC_WALL_FN_CAP;
    m = k * 1e3 / (wc * wc);
  }
  w->k = k; w->m = m; w->fn = fn; w->zeta = 0.02; w->nx = nx; w->ny = ny; w->t = t; w->tRoot = ts[0];
  w->hWall = hWall; w->a = a; w->bEff = bEff; w->base = base; w->cx = cx; w->cy = cy; w->I = I;
  return 1;
}

/* wallAt (src/wall.js:11-34): Hb holds MC_BINS engagement heights (the engine's map after
   fillEngagement), (ux, uy) the feed direction.  Returns 1 and fills the wall fields of *w, or
   0 (null: the workpiece is rigid for this sub-step). */
int or_wall_at(const or_stock *s, double px, double py, double pz, double R, const float *Hb, double ux, double uy,
               double zLoad, double E, double rho, double kTool, mc_wall_out *w) {
  const double vx = -uy, vy = ux;
  double sx = 0, sy = 0, L, n0, cx, cy, bt = 0, bnx = 0, bny = 0, d;
  int b, have = 0;
  for (b = 0; b < MC_BINS / 2; b++) {
    double phi, sn, cs;
    if ((double)Hb[b] <= zLoad) continue;
    phi = (b + 0.5) / MC_BINS * 2 * PI;
    sn = or_sin(phi);
    cs = or_cos(phi);
    sx += sn * ux + cs * vx;
    sy += sn * uy + cs * vy;
  }
  L = or_hypot(sx, sy);
  if (L < 1e-9) return 0;
  n0 = or_atan2(sy, sx);
  cx = px + R * sx / L;
  cy = py + R * sy / L;
  for (d = -40; d <= 40; d += 5) {
    const double a = n0 + d * PI / 180, ca = or_cos(a), sa = or_sin(a);
    const double t = thickness(s, cx, cy, ca, sa, zLoad);
    if (t > 0 && (!have || t < bt)) { have = 1; bt = t; bnx = ca; bny = sa; }
  }
  if (!have) return 0;
  return measure(s, pz, cx, cy, R, bnx, bny, bt, zLoad, E, rho, kTool, w);
}
