/*
 * oracle/simloop.c — TEST INFRASTRUCTURE ONLY (the oracle).  CPU restatement of the
 * simulator's sub-step loop with a rigid workpiece and a fixed flank wear, serially in the
 * reference's order (src/sim.js:380-445):
 *
 *   fillEngagement     src/sim.js:301-319  (or_fill_engagement, engage.c)
 *   fillEndEngagement  src/sim.js:324-341  (below)
 *   axial depth, end cut, the cut decision                 src/sim.js:388-401
 *   setEngagement / setEndEngagement, the wall, setWork    src/sim.js:403-408
 *     (workpieceFlex: wallAt at the load height tipZ + ap / 2, oracle/wall.c; else null)
 *   the run's step count (the frame budget's share)        src/sim.js:358-362,409-414
 *   the frame's end: takeStats() and chatterIndex()        src/sim.js:450,645-663
 *   the engagement arc it shows (immersion, :343-348)      src/sim.js:421-423
 *   the decay out of the cut, a decayed wall dropped       src/sim.js:425-429
 *   deflection, the commanded and the actual stamp         src/sim.js:433-440
 *   the wall-error mark (markWall) while a wall is held    src/sim.js:443
 *
 * The engine is oracle/cengine.c's object, the stocks oracle/stock.c's.  The tool is flat,
 * ball or bull-nose; the end-face shape zAt is the simulator's (src/sim.js:216-218: the tip
 * table when the tool has one, the sphere for a ball, 0 for a flat end), and a bull-nose
 * tool's slices follow cutterRadiusAt (src/assembly.js:23-30) as tool.rAt
 * (src/sim.js:190-191, src/cutting.js:113-119).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define PI 3.141592653589793

typedef struct { int ball; double R; const float *lut; int lutn; double rc; const double *er; int en; double edz; } Shape;

static double zat(const Shape *t, double r) {
  if (t->lut) {
    double i = or_js_round(r / MC_LUT_DR);
    if (i > t->lutn - 1) i = t->lutn - 1;
    return t->lut[(int)i];
  }
  return t->ball ? t->R - sqrt(fmax(0.0, t->R * t->R - r * r)) : 0.0;
}

/* cutterRadiusAt (src/assembly.js:23-30) for a bull-nose corner rc */
static double bull_rat(double R, double rc, double z) {
  if (rc > 0 && z < rc) {
    const double q = rc - z, v = rc * rc - q * q;
    return R - rc + sqrt(v > 0 ? v : 0);
  }
  return R;
}

/* cutterRadiusAt of an imported edge: sampleEdge (src/assembly.js:31-35), the radii en
   samples edz apart, linear between them, the end samples held beyond the table */
static double edge_rat(const Shape *t, double z) {
  const double f = z / t->edz;
  double i = floor(f), u;
  if (i < 0) i = 0;
  if (i > t->en - 2) i = t->en - 2;
  u = f - i;
  if (u < 0) u = 0;
  if (u > 1) u = 1;
  return t->er[(int)i] * (1 - u) + t->er[(int)i + 1] * u;
}

static double profile_rat(const Shape *t, double z) { return t->er ? edge_rat(t, z) : bull_rat(t->R, t->rc, z); }

/* Simulator.fillEndEngagement (src/sim.js:324-341): Eb[b * 6 + j]; returns dmax, *rOut */
static double fill_end(const or_stock *s, double cx, double cy, double tipZ, double ux, double uy,
                       const Shape *t, float *Eb, double *rOut_out) {
  const double R = t->R;
  int32_t nx, ny;
  double hx, hy, rOut, dmax = 0, vx = -uy, vy = ux;
  int b, j;
  or_stock_dims(s, &nx, &ny, &hx, &hy);
  rOut = R - 0.5 * (hx > hy ? hx : hy);
  for (b = 0; b < MC_END_BINS; b++) {
    const double phi = (b + 0.5) / MC_END_BINS * 2 * PI, sn = or_sin(phi), cs = or_cos(phi);
    const double ex = sn * ux + cs * vx, ey = sn * uy + cs * vy;
    for (j = 0; j < MC_END_RINGS; j++) {
      const double r = (j + 0.5) / MC_END_RINGS * rOut;
      const double zEdge = tipZ + zat(t, r);
      double d = or_stock_sample(s, cx + r * ex, cy + r * ey) - zEdge;
      if (!(d > 0)) d = 0; /* Math.max(0, d): d is never NaN here */
      Eb[b * MC_END_RINGS + j] = (float)d;
      if (d > dmax) dmax = d;
    }
  }
  *rOut_out = rOut;
  return dmax;
}

double or_fill_end_engagement(const or_stock *s, double cx, double cy, double tipZ, double ux, double uy,
                              double R, int ball, float *Eb, double *rOut_out) {
  const Shape t = {ball, R, NULL, 0, 0, NULL, 0, 0};
  return fill_end(s, cx, cy, tipZ, ux, uy, &t, Eb, rOut_out);
}

/* setEndEngagement's rings (src/cutting.js:133-142) for the simulator's zAt */
static void end_rings(double rOut, const Shape *t, double *g) {
  const double dr = rOut / MC_END_RINGS;
  int j;
  for (j = 0; j < MC_END_RINGS; j++) {
    const double r = (j + 0.5) * dr;
    const double a = r + 0.25 * dr < rOut ? r + 0.25 * dr : rOut;
    const double b = r - 0.25 * dr > 0 ? r - 0.25 * dr : 0;
    const double slope = (zat(t, a) - zat(t, b)) / (0.5 * dr);
    double c = 1 / sqrt(1 + slope * slope);
    if (isnan(c)) c = NAN; else if (c < 0.22) c = 0.22;
    g[j] = r; g[MC_END_RINGS + j] = c; g[2 * MC_END_RINGS + j] = zat(t, r);
    g[3 * MC_END_RINGS + j] = dr / c;
  }
}

int or_sim_run(or_stock *act, or_stock *ide, or_engine *e, const mc_substep *sub, int n, double R,
               double fluteLen, double rw, double share, int type, const float *lut, int lutn, double cornerR,
               mc_substep_out *out, double dynBudget, double fz0, mc_frame_out *frames, int nframes, int flex,
               double E, double rho, double kTool, mc_wall_out *walls, const double *edge_r, int edge_n,
               double edge_dz, double *budgetLeft) {
  const int ball = type == MC_TOOL_BALL, prof = type == MC_TOOL_PROFILE, bull = type == MC_TOOL_BULL || prof;
  const Shape shape = {ball, R, bull ? lut : NULL, bull ? lutn : 0,
                       type == MC_TOOL_BULL ? (cornerR < R ? cornerR : R) : 0, prof ? edge_r : NULL,
                       prof ? edge_n : 0, edge_dz};
  const float *slut = bull ? lut : NULL;
  const int stype = ball ? MC_TOOL_BALL : MC_TOOL_FLAT;
  float Hb[MC_BINS], Eb[MC_END_BINS * MC_END_RINGS];
  double fzLast = fz0, dynLeft = budgetLeft ? *budgetLeft : dynBudget;   /* a call may begin inside a frame */
  int k, frame = 0, held = 0;   /* held: this.wall is not null */
  memset(Hb, 0, sizeof(Hb));
  memset(Eb, 0, sizeof(Eb));
  for (k = 0; k < n; k++) {
    const mc_substep *s = &sub[k];
    mc_engine_state st;
    double hmax = -INFINITY, ap, dmax = 0, rOut = 0, dx, dy, errVal, vol, maxF = 0;
    int haveEnd = s->fza > 0, endCut, cut, nsteps = 0, arc;
    double phiSt, phiEx;
    mc_wall_out wo;
    memset(&wo, 0, sizeof(wo));
    if (s->frame & MC_FRAME_FIRST) dynLeft = dynBudget;
    if (!(s->dxy < 1e-6)) hmax = or_fill_engagement(act, s->x, s->y, s->ux, s->uy, R, Hb);
    ap = hmax - s->z > 0 ? hmax - s->z : 0;
    if (fluteLen < ap) ap = fluteLen;
    if (haveEnd) dmax = fill_end(act, s->x, s->y, s->z, s->ux, s->uy, &shape, Eb, &rOut);
    endCut = haveEnd && dmax > 2e-4;
    cut = (ap > 0.01 || endCut) && !s->rapid;
    /* the engagement arc the simulator shows (src/sim.js:421-423) */
    phiSt = 0; phiEx = 0; arc = 0;
    if (cut && ap > 0.01) arc = or_immersion(Hb, s->z, &phiSt, &phiEx);
    else if (cut && endCut) { phiEx = PI; arc = 1; }
    or_engine_get_state(e, &st);
    if (cut) {
      double g[4 * MC_END_RINGS], want;
      if (bull) {
        /* setEngagement's rAt slices (src/cutting.js:112-119): a bull-nose corner or an
           imported edge */
        double sr[MC_MAX_SLICES], sk[MC_MAX_SLICES];
        const double dz = or_engine_dz(e);
        const int ms = or_engine_max_slices(e);
        int k, nz = ap <= 0 ? 0 : (int)fmin(ms, ceil(ap / dz));
        for (k = 0; k < nz; k++) {
          const double b = dz < ap - k * dz ? dz : ap - k * dz, zr = k * dz + 0.5 * b;
          const double r0 = profile_rat(&shape, zr);
          const double lo = zr - 0.05 > 0 ? zr - 0.05 : 0;
          const double dr = (profile_rat(&shape, zr + 0.05) - profile_rat(&shape, lo)) / (zr < 0.05 ? 0.05 + zr : 0.1);
          sr[k] = r0 > 1e-3 ? r0 : 1e-3;
          sk[k] = 1 / sqrt(1 + dr * dr);
        }
        or_engine_set_engagement(e, s->ux, s->uy, s->fzEff, s->z, ap, sr, sk, NULL);
      } else {
        or_engine_set_engagement(e, s->ux, s->uy, s->fzEff, s->z, ap, NULL, NULL, NULL);
      }
      end_rings(haveEnd ? rOut : 0, &shape, g);
      or_engine_set_end_engagement(e, endCut ? s->fza : 0, haveEnd ? rOut : 0, g);
      held = flex && ap > 0.01 &&
             or_wall_at(act, s->x, s->y, s->z, R, Hb, s->ux, s->uy, s->z + ap / 2, E, rho, kTool, &wo);
      wo.found = held;
      if (held) {
        const mc_wall wl = {wo.k, wo.m, wo.zeta, wo.nx, wo.ny};
        or_engine_set_work(e, &wl);
      } else {
        or_engine_set_work(e, NULL);
      }
      or_engine_set_maps(e, Hb, Eb);
      want = or_js_round(s->dtm / or_engine_dt(e));
      if (want < 1) want = 1;
      if (dynBudget > 0) {
        /* cost = tool.N * max(1, e.nz + (endCut ? END_RINGS : 0)); share = max(spt / 4, dynLeft
           / cost / max(1, nSubEst - guard + 1)) (src/sim.js:410-411) */
        mc_engine_state es;
        double cost, sh, den;
        or_engine_get_state(e, &es);
        cost = (double)or_engine_flutes(e) * (es.nz + (endCut ? MC_END_RINGS : 0) > 1 ? es.nz + (endCut ? MC_END_RINGS : 0) : 1);
        den = s->nSubEst - s->guard + 1;
        if (!(den > 1)) den = 1;
        sh = dynLeft / cost / den;
        if (!(sh > or_engine_spt(e) / 4.0)) sh = or_engine_spt(e) / 4.0;
        nsteps = (int)(want < floor(sh) ? want : floor(sh));
        if (nsteps < 1) nsteps = 1;
        dynLeft -= nsteps * cost;
      } else {
        nsteps = (int)(want < floor(share) ? want : floor(share));
        if (nsteps < 1) nsteps = 1;
      }
      or_engine_run(e, nsteps);
      maxF = or_engine_run_max_f(e);
      fzLast = s->fzEff;
    } else {
      st.x *= 0.5; st.y *= 0.5; st.vx *= 0.5; st.vy *= 0.5; st.w *= 0.5; st.wv *= 0.5;
      or_engine_set_state(e, &st);
      if (st.wk != 0 && fabs(st.w) < 1e-7) { held = 0; or_engine_set_work(e, NULL); }
    }
    or_engine_get_state(e, &st);
    dx = (st.x - st.w * st.wnx) * 1e3;
    dy = (st.y - st.w * st.wny) * 1e3;
    errVal = or_hypot(st.x, st.y) * 1e6;
    or_stock_stamp_ex(ide, s->x, s->y, s->z, R, stype, 0, s->line, NULL, slut, shape.lutn);
    {
      const double mark[3] = {s->ux, s->uy, fzLast * or_engine_flutes(e)};
      vol = or_stock_stamp_ex(act, s->x + dx, s->y + dy, s->z + (ball ? rw : 0), R - rw, stype, errVal, s->line, mark,
                              slut, shape.lutn);
    }
    if (held && st.wk != 0) {
      const double v = st.runMeanW * 1e6;
      wo.marked = 1; wo.markVal = v > 0 ? v : 0; wo.wnx = st.wnx; wo.wny = st.wny;
      or_stock_mark_wall(act, s->x + dx, s->y + dy, R, st.wnx, st.wny, wo.markVal);
    }
    if (walls) walls[k] = wo;
    if (out) {
      out[k].dx = dx; out[k].dy = dy; out[k].ap = ap; out[k].errVal = errVal; out[k].vol = vol;
      out[k].maxF = maxF; out[k].nsteps = nsteps; out[k].cut = cut;
      out[k].phiSt = arc ? phiSt : 0; out[k].phiEx = arc ? phiEx : 0; out[k].arc = arc; out[k].pad = 0;
    }
    /* Simulator.collect at the frame's end: takeStats(), then chatterIndex() */
    if ((s->frame & MC_FRAME_LAST) && frames && frame < nframes) {
      double ci[3];
      or_engine_stats(e, &frames[frame].stats, 1);
      or_engine_chatter_index(e, ci);
      frames[frame].index = ci[0]; frames[frame].amp = ci[1]; frames[frame].once = ci[2];
      frames[frame].end = k + 1; frames[frame].pad = 0;
      frame++;
    }
  }
  if (budgetLeft) *budgetLeft = dynLeft;
  return MC_OK;
}
