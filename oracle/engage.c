/*
 * oracle/engage.c — TEST INFRASTRUCTURE ONLY (the oracle).  Restates the simulator's
 * engagement sampling on the stock:
 *
 *   fillEngagement  src/sim.js:301-319  (stock heights on the cutter periphery, one per
 *                                        angle bin, f32 like engine.Hb)
 *   immersion       src/sim.js:343-348  (engagement angles at the bottom slice)
 *   ap              src/sim.js:388      (axial depth from the highest bin)
 *
 * FP64 arithmetic, V8's sin / cos (fdlibm), the stock's bilinear sample.
 */
#include <math.h>

#include "oracle.h"

#define PI 3.141592653589793

/* returns hmax; Hb receives MC_BINS f32 heights (-inf behind the cutter or off the stock) */
double or_fill_engagement(const or_stock *s, double cx, double cy, double ux, double uy, double R,
                          float *Hb) {
  int32_t nx, ny;
  double hx, hy, cell, r, ax, ay, vx = -uy, vy = ux, hmax = -INFINITY;
  int b;
  or_stock_dims(s, &nx, &ny, &hx, &hy);
  cell = hx > hy ? hx : hy;
  r = R - 0.5 * cell;
  ax = cx + ux * cell;
  ay = cy + uy * cell;
  for (b = 0; b < MC_BINS; b++) {
    const double phi = (b + 0.5) / MC_BINS * 2 * PI;
    const double sn = or_sin(phi), cs = or_cos(phi);
    const double h = or_stock_sample(s, ax + r * (sn * ux + cs * vx), ay + r * (sn * uy + cs * vy));
    Hb[b] = sn > 0.02 ? (float)h : -INFINITY;
    if (Hb[b] > hmax) hmax = Hb[b];
  }
  return hmax;
}

/* returns 1 and the arc [st, ex] when some leading bin stands above tipZ + 0.1, else 0 */
int or_immersion(const float *Hb, double tipZ, double *st, double *ex) {
  const double z = tipZ + 0.1;
  int b, found = 0;
  double a = 0, e = 0;
  for (b = 0; b < MC_BINS / 2; b++)
    if (Hb[b] > z) {
      const double phi = (b + 0.5) / MC_BINS * 2 * PI;
      if (!found) a = phi;
      found = 1;
      e = phi;
    }
  if (!found) return 0;
  *st = a - PI / MC_BINS > 0 ? a - PI / MC_BINS : 0;
  *ex = e + PI / MC_BINS < PI ? e + PI / MC_BINS : PI;
  return 1;
}
