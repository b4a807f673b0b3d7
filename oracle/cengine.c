/*
 * oracle/cengine.c — TEST INFRASTRUCTURE ONLY (the oracle).  CPU restatement, in FP64 with
 * the reference's float32 store points and summation order, of the whole CuttingEngine
 * object as a caller keeps it across calls (src/cutting.js:21-326):
 *
 *   constructor            src/cutting.js:22-35
 *   configure              src/cutting.js:39-82
 *   setWear / setWork      src/cutting.js:85-96
 *   setEngagement          src/cutting.js:98-124  (rAt slices taken from the caller)
 *   setEndEngagement       src/cutting.js:130-143 (rings taken from the caller, or zAt = 0)
 *   run                    src/cutting.js:146-296 (side and end edges, wall mode, stats,
 *                                                  history, tooth and wall samples, chip
 *                                                  events)
 *   chatterIndex           src/cutting.js:299-318
 *   takeStats / newStats   src/cutting.js:320-326
 *
 * A JavaScript value read out of a typed array's range is undefined; the comparisons it
 * meets here (undefined <= z, undefined < fza) are false, which NaN reproduces.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define BINS 180
#define EBINS MC_END_BINS
#define ERINGS MC_END_RINGS
#define HN MC_ENGINE_HIST
#define RN MC_ENGINE_RING
#define H_REF 0.05
#define H_MIN 0.002
#define JS_PI 3.141592653589793

enum { S_N, S_SUMF, S_MAXF, S_SUMFX, S_SUMFY, S_SUMFZ, S_MAXFZ, S_SUMT, S_MAXT, S_SUMX, S_SUMY, S_MINX,
       S_MAXX, S_MINY, S_MAXY, S_SUMZC, S_WZC, S_SUMW, S_MAXW, S_SUMHB, S_SUMB, S_SUMRB, S_ENGT, S_COUNT };

struct or_engine {
  int configured, engaged;
  /* configure */
  mc_tool tool;
  mc_material mat;
  int nonlinear, N, spt, spr, decim, piStep, maxSlices, nmx, nmy;
  double R, tanB, S, omega, Ttooth, dt, dz;
  double dR[MC_MAX_FLUTES];
  double kx[MC_MAX_MODES], mx[MC_MAX_MODES], cx[MC_MAX_MODES], qx[MC_MAX_MODES], vx[MC_MAX_MODES];
  double ky[MC_MAX_MODES], my[MC_MAX_MODES], cy[MC_MAX_MODES], qy[MC_MAX_MODES], vy[MC_MAX_MODES];
  float *surf;
  int pr;
  double x, y, vxs, vys, x0, y0, vx0, vy0, t;
  int64_t stepCount;
  /* setWear: undefined until called */
  int wear;
  double Kte, Kre;
  /* setWork */
  double runMeanW, wk, wm, wc, w, wv, wnx, wny;
  /* setEngagement */
  double ux, uy, ex, ey, fz, tipZ, db;
  int nz;
  double sB[MC_MAX_SLICES], sZ[MC_MAX_SLICES], sLag[MC_MAX_SLICES], sR[MC_MAX_SLICES], sKap[MC_MAX_SLICES];
  /* setEndEngagement */
  double fza, endR, ringR[ERINGS], ringCos[ERINGS], ringZ[ERINGS], ringDs[ERINGS];
  float Hb[BINS], Eb[EBINS * ERINGS];
  /* per tooth */
  int toothEngaged[MC_MAX_FLUTES];
  double toothVol[MC_MAX_FLUTES], toothPhi[MC_MAX_FLUTES], toothZ[MC_MAX_FLUTES];
  /* stats and rings */
  double st[S_COUNT];
  float hFx[HN], hFy[HN], hFz[HN], hx[HN], hy[HN];
  int64_t hist_i, hist_count;
  float tsx[RN], tsy[RN], wsx[RN], wsy[RN], wpx[RN], wpy[RN];
  int64_t ts_i, ts_count, ws_i, ws_count, wsp_i, wsp_count;
  /* chip events of the last run */
  mc_chip_event *ev;
  int64_t nev, ev_cap;
  double runMaxF; /* max |(Fx, Fy)| over the last run */
};

static double jmax(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == 0 && b == 0) return signbit(a) ? b : a;
  return a > b ? a : b;
}
static double jmin(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == 0 && b == 0) return signbit(a) ? a : b;
  return a < b ? a : b;
}
static int truthy(double v) { return v != 0 && !isnan(v); }

static void new_stats(double *st) {
  memset(st, 0, sizeof(double) * S_COUNT);
  st[S_MINX] = INFINITY; st[S_MAXX] = -INFINITY;
  st[S_MINY] = INFINITY; st[S_MAXY] = -INFINITY;
}

or_engine *or_engine_create(void) {
  or_engine *e = (or_engine *)calloc(1, sizeof(or_engine));
  if (!e) return NULL;
  new_stats(e->st);
  return e;
}

void or_engine_destroy(or_engine *e) {
  if (!e) return;
  free(e->surf);
  free(e->ev);
  free(e);
}

static void set_work_null(or_engine *e) {
  e->runMeanW = 0; e->wk = 0; e->wm = 1; e->wc = 0; e->w = 0; e->wv = 0; e->wnx = 1; e->wny = 0;
}

/* configure(tool, modal, mat, S, fnMax), src/cutting.js:39-82 */
int or_engine_configure(or_engine *e, const mc_cell *c, double fnMax, mc_engine_info *info) {
  int i;
  double fTop = -INFINITY, rem;
  const double TWO_PI = 2 * JS_PI;
  if (!e || !c || c->tool.N < 1 || c->tool.N > MC_MAX_FLUTES || c->nmx < 1 || c->nmx > MC_MAX_MODES ||
      c->nmy < 1 || c->nmy > MC_MAX_MODES)
    return MC_ERR_ARG;
  e->tool = c->tool; e->mat = c->mat; e->nonlinear = c->nonlinear;
  e->N = c->tool.N; e->R = c->tool.R;
  e->tanB = or_tan(c->tool.helixDeg * JS_PI / 180);
  for (i = 0; i < e->N; i++) e->dR[i] = c->tool.runout * or_cos(2 * JS_PI * i / e->N - c->tool.runoutPhase);
  e->S = jmax(100, c->S);
  e->omega = e->S * 2 * JS_PI / 60;
  e->Ttooth = 60 / (e->S * e->N);
  for (i = 0; i < c->nmx; i++) fTop = jmax(fTop, c->mx[i].fn);
  for (i = 0; i < c->nmy; i++) fTop = jmax(fTop, c->my[i].fn);
  e->spt = (int)jmax(48, ceil(e->Ttooth * jmax(fTop, fnMax) * 24));
  e->dt = e->Ttooth / e->spt;
  e->nmx = c->nmx; e->nmy = c->nmy;
  for (i = 0; i < c->nmx; i++) {
    double w = 2 * JS_PI * c->mx[i].fn;
    e->kx[i] = c->mx[i].k * 1e3; e->mx[i] = e->kx[i] / (w * w);
    e->cx[i] = 2 * c->mx[i].zeta * sqrt(e->kx[i] * e->mx[i]);
    e->qx[i] = 0; e->vx[i] = 0;
  }
  for (i = 0; i < c->nmy; i++) {
    double w = 2 * JS_PI * c->my[i].fn;
    e->ky[i] = c->my[i].k * 1e3; e->my[i] = e->ky[i] / (w * w);
    e->cy[i] = 2 * c->my[i].zeta * sqrt(e->ky[i] * e->my[i]);
    e->qy[i] = 0; e->vy[i] = 0;
  }
  e->x0 = 0; e->y0 = 0; e->vx0 = 0; e->vy0 = 0;
  e->dz = jmax(0.25, c->tool.fluteLen / 48);
  e->maxSlices = (int)ceil(c->tool.fluteLen / e->dz);
  if (e->maxSlices > MC_MAX_SLICES) return MC_ERR_LIMIT;
  e->spr = e->N * e->spt;
  free(e->surf);
  e->surf = (float *)calloc((size_t)e->maxSlices * e->spr, sizeof(float));
  if (!e->surf) return MC_ERR_NOMEM;
  e->pr = 0;
  e->x = 0; e->y = 0; e->vxs = 0; e->vys = 0; e->t = 0;
  set_work_null(e);
  e->fza = 0;
  memset(e->toothEngaged, 0, sizeof(e->toothEngaged));
  memset(e->toothVol, 0, sizeof(e->toothVol));
  memset(e->toothPhi, 0, sizeof(e->toothPhi));
  memset(e->toothZ, 0, sizeof(e->toothZ));
  e->decim = (int)jmax(1, or_js_round(e->spt / 40.0));
  rem = fmod(JS_PI, TWO_PI / e->N);
  e->piStep = (int)fmod(or_js_round((rem / (TWO_PI / e->N)) * e->spt), e->spt);
  e->ws_i = 0; e->ws_count = 0; e->wsp_i = 0; e->wsp_count = 0;
  e->stepCount = 0;
  e->configured = 1;
  e->engaged = 0;
  if (info) {
    info->dt = e->dt; info->omega = e->omega; info->Ttooth = e->Ttooth; info->dz = e->dz; info->S = e->S;
    info->stepsPerTooth = e->spt; info->stepsPerRev = e->spr; info->decim = e->decim;
    info->piStep = e->piStep; info->maxSlices = e->maxSlices; info->nmx = e->nmx; info->nmy = e->nmy;
    info->pad = 0;
  }
  return MC_OK;
}

void or_engine_set_wear(or_engine *e, double Kte, double Kre) { e->wear = 1; e->Kte = Kte; e->Kre = Kre; }

/* setWork(wall), src/cutting.js:88-96 */
void or_engine_set_work(or_engine *e, const mc_wall *w) {
  if (!w) { set_work_null(e); return; }
  if (truthy(e->wk) && w->nx * e->wnx + w->ny * e->wny < 0.8) { e->w = 0; e->wv = 0; }
  e->wk = w->k * 1e3; e->wm = w->m; e->wc = 2 * w->zeta * sqrt(e->wk * w->m);
  e->wnx = w->nx; e->wny = w->ny;
}

/* setEngagement(u, fz, tipZ, ap), src/cutting.js:98-124 */
int or_engine_set_engagement(or_engine *e, double ux, double uy, double fz, double tipZ, double ap,
                             const double *sliceR, const double *sliceKap, int32_t *nz_out) {
  int k, nz;
  if (!e->configured) return MC_ERR_ARG;
  e->ux = ux; e->uy = uy; e->ex = -uy; e->ey = ux;
  e->fz = fz; e->tipZ = tipZ;
  nz = ap <= 0 ? 0 : (int)jmin(e->maxSlices, ceil(ap / e->dz));
  e->nz = nz;
  e->db = e->dz;
  for (k = 0; k < nz; k++) {
    double zr, r = e->R, kap = 1;
    e->sB[k] = jmin(e->dz, ap - k * e->dz);
    zr = k * e->dz + 0.5 * e->sB[k];
    e->sZ[k] = tipZ + zr;
    if (sliceR) { r = sliceR[k]; kap = sliceKap[k]; }
    else if (e->tool.type == MC_TOOL_BALL && zr < e->R) {
      double q = e->R - zr;
      r = sqrt(jmax(1e-6, e->R * e->R - q * q));
      kap = r / e->R;
    }
    e->sR[k] = r; e->sKap[k] = kap;
    e->sLag[k] = zr * e->tanB / e->R;
  }
  e->engaged = 1;
  if (nz_out) *nz_out = nz;
  return MC_OK;
}

/* setEndEngagement(fza, rOut, zAt), src/cutting.js:130-143; rings = NULL is zAt = () => 0 */
void or_engine_set_end_engagement(or_engine *e, double fza, double rOut, const double *rings) {
  int j;
  const double dr = rOut / ERINGS;
  e->fza = fza > 0 ? fza : 0;
  e->endR = rOut;
  for (j = 0; j < ERINGS; j++) {
    if (rings) {
      e->ringR[j] = rings[j]; e->ringCos[j] = rings[ERINGS + j];
      e->ringZ[j] = rings[2 * ERINGS + j]; e->ringDs[j] = rings[3 * ERINGS + j];
    } else {
      const double r = (j + 0.5) * dr;
      const double slope = (0.0 - 0.0) / (0.5 * dr);
      const double c = jmax(0.22, 1 / sqrt(1 + slope * slope));
      e->ringR[j] = r; e->ringCos[j] = c; e->ringDs[j] = dr / c; e->ringZ[j] = 0;
    }
  }
}

void or_engine_set_maps(or_engine *e, const float *Hb, const float *Eb) {
  if (Hb) memcpy(e->Hb, Hb, sizeof(e->Hb));
  if (Eb) memcpy(e->Eb, Eb, sizeof(e->Eb));
}

void or_engine_get_state(const or_engine *e, mc_engine_state *s) {
  memset(s, 0, sizeof(*s));
  s->x = e->x; s->y = e->y; s->vx = e->vxs; s->vy = e->vys; s->w = e->w; s->wv = e->wv; s->t = e->t;
  s->runMeanW = e->runMeanW; s->wk = e->wk; s->wnx = e->wnx; s->wny = e->wny;
  s->stepCount = e->stepCount; s->pr = e->pr; s->decim = e->decim; s->nz = e->nz;
  s->hist_i = e->hist_i; s->hist_count = e->hist_count; s->ts_i = e->ts_i; s->ts_count = e->ts_count;
  s->ws_i = e->ws_i; s->ws_count = e->ws_count; s->wsp_i = e->wsp_i; s->wsp_count = e->wsp_count;
}

void or_engine_set_state(or_engine *e, const mc_engine_state *s) {
  e->x = s->x; e->y = s->y; e->vxs = s->vx; e->vys = s->vy; e->w = s->w; e->wv = s->wv; e->t = s->t;
  e->decim = s->decim; e->hist_i = s->hist_i;
}

/* run(nSteps), src/cutting.js:146-296 */
int or_engine_run(or_engine *e, int64_t nSteps) {
  const int N = e->N, spt = e->spt, spr = e->spr, nz = e->nz, MS = e->maxSlices;
  const double ux = e->ux, uy = e->uy, vx = e->ex, vy = e->ey, fz = e->fz, dt = e->dt, fza = e->fza;
  const double Ktc = e->mat.Ktc, Krc = e->mat.Krc, Kac = e->mat.Kac, Kae = e->mat.Kae;
  const double Kte = e->wear ? e->Kte : e->mat.Kte, Kre = e->wear ? e->Kre : e->mat.Kre;
  const int nl = e->nonlinear;
  const double mc = nl ? e->mat.mc : 0;
  const double pdK = nl ? e->mat.Ci * e->db * 1e-3 / (e->omega * 1e-3) : 0;
  const double wnx = e->wnx, wny = e->wny;
  const double TWO_PI = 2 * JS_PI, binScale = BINS / TWO_PI, pitch = TWO_PI / N;
  const double dphi = e->omega * dt;
  const double hMax = jmax(0.1, 3 * fz);
  const double wk = e->wk, wm = e->wm, wc = e->wc;
  double *st = e->st;
  double x, y, vxs, vys, t, w, wv, wSum = 0;
  int pr;
  int64_t s, cap;
  if (!e->configured || !e->engaged || nSteps < 0) return MC_ERR_ARG;
  cap = (int64_t)N * (nSteps / 2 + 2);
  if (cap > e->ev_cap) {
    free(e->ev);
    e->ev = (mc_chip_event *)malloc(sizeof(mc_chip_event) * (size_t)cap);
    if (!e->ev) { e->ev_cap = 0; return MC_ERR_NOMEM; }
    e->ev_cap = cap;
  }
  e->nev = 0;
  e->runMaxF = 0;
  e->qx[0] += e->x - e->x0; e->qy[0] += e->y - e->y0;
  e->vx[0] += e->vxs - e->vx0; e->vy[0] += e->vys - e->vy0;
  x = e->x; y = e->y; vxs = e->vxs; vys = e->vys; t = e->t; pr = e->pr;
  w = e->w; wv = e->wv;
  for (s = 0; s < nSteps; s++) {
    const double theta = pr * dphi;
    const double rx = x - w * wnx, ry = y - w * wny;
    const double xm = rx * 1e3, ym = ry * 1e3;
    double Fx = 0, Fy = 0, Fz = 0, T = 0, Fr2z = 0, Fsum = 0, Fxy;
    const double rvx = vxs - wv * wnx, rvy = vys - wv * wny;
    int i, j, k, ph;
    for (i = 0; i < N; i++) {
      int engaged = 0;
      double phs = 0, zsum = 0;
      const double dRi = e->dR[i];
      int pi = pr + i * spt;
      if (pi >= spr) pi -= spr;
      for (k = 0; k < nz; k++) {
        double phi = fmod(theta + i * pitch - e->sLag[k], TWO_PI);
        double sn, cs, nxg, nyg, rn, prev, feedH, h, se, bk, Ft, Fr, Fa, txg, tyg, fr, hb;
        int idx, bin;
        if (phi < 0) phi += TWO_PI;
        sn = or_sin(phi); cs = or_cos(phi);
        nxg = sn * ux + cs * vx; nyg = sn * uy + cs * vy;
        rn = xm * nxg + ym * nyg + dRi;
        idx = pi * MS + k;
        prev = e->surf[idx];
        feedH = fz * sn * e->sKap[k];
        bin = (int)(phi * binScale);
        hb = bin < BINS ? (double)e->Hb[bin] : NAN;
        if (hb <= e->sZ[k] || sn <= 0) { e->surf[idx] = (float)rn; continue; }
        h = feedH + rn - prev;
        if (h <= 0) { e->surf[idx] = (float)(prev - feedH); continue; }
        e->surf[idx] = (float)rn;
        if (h > hMax) h = hMax;
        se = truthy(mc) ? or_pow(H_REF / (h > H_MIN ? h : H_MIN), mc) : 1;
        bk = e->sB[k];
        Ft = Ktc * se * h * bk + Kte * bk;
        Fr = Krc * se * h * bk + Kre * bk;
        Fa = Kac * se * h * bk + Kae * bk;
        if (truthy(pdK)) Fr += pdK * (bk / e->db) / e->sR[k] * (rvx * nxg + rvy * nyg);
        txg = cs * ux - sn * vx; tyg = cs * uy - sn * vy;
        Fx += -Ft * txg - Fr * nxg;
        Fy += -Ft * tyg - Fr * nyg;
        Fz -= Fa;
        T += Ft * e->sR[k];
        fr = or_hypot(Ft, Fr);
        Fsum += fr; Fr2z += fr * (e->sZ[k] - e->tipZ);
        engaged = 1; phs = phi; zsum = e->sZ[k];
        st[S_SUMHB] += h * bk; st[S_SUMB] += bk; st[S_SUMRB] += e->sR[k] * h * bk;
        e->toothVol[i] += h * bk * e->sR[k] * dphi;
      }
      if (truthy(fza)) {
        double phi = fmod(theta + i * pitch, TWO_PI), sn, cs, txg, tyg;
        int rb;
        if (phi < 0) phi += TWO_PI;
        rb = (int)(phi / TWO_PI * EBINS);
        sn = or_sin(phi); cs = or_cos(phi);
        txg = cs * ux - sn * vx; tyg = cs * uy - sn * vy;
        for (j = 0; j < ERINGS; j++) {
          const double depth = rb < EBINS ? (double)e->Eb[rb * ERINGS + j] : NAN;
          double c, ds, h, se, Ft, Fn;
          if (depth <= 2e-4) continue;
          c = e->ringCos[j]; ds = e->ringDs[j];
          h = (depth < fza ? depth : fza) * c;
          se = truthy(mc) ? or_pow(H_REF / (h > H_MIN ? h : H_MIN), mc) : 1;
          Ft = Ktc * se * h * ds + Kte * ds;
          Fn = Krc * se * h * ds + Kre * ds;
          Fx += -Ft * txg; Fy += -Ft * tyg;
          Fz -= Fn * c;
          T += Ft * e->ringR[j];
          Fsum += Ft; Fr2z += Ft * e->ringZ[j];
          engaged = 1; phs = phi; zsum = e->tipZ + e->ringZ[j];
          st[S_SUMHB] += h * ds; st[S_SUMB] += ds; st[S_SUMRB] += e->ringR[j] * h * ds;
          e->toothVol[i] += h * ds * e->ringR[j] * dphi;
        }
      }
      if (engaged) { e->toothPhi[i] = phs; e->toothZ[i] = zsum; }
      else if (e->toothEngaged[i]) {
        if (e->toothVol[i] > 0) {
          mc_chip_event *c = &e->ev[e->nev++];
          c->vol = e->toothVol[i]; c->phi = e->toothPhi[i]; c->z = e->toothZ[i];
          c->ux = ux; c->uy = uy; c->vx = vx; c->vy = vy;
        }
        e->toothVol[i] = 0;
      }
      e->toothEngaged[i] = engaged;
      st[S_ENGT] += engaged;
    }
    x = 0; y = 0; vxs = 0; vys = 0;
    for (j = 0; j < e->nmx; j++) {
      e->vx[j] += (Fx - e->cx[j] * e->vx[j] - e->kx[j] * e->qx[j]) / e->mx[j] * dt;
      e->qx[j] += e->vx[j] * dt; x += e->qx[j]; vxs += e->vx[j];
    }
    for (j = 0; j < e->nmy; j++) {
      e->vy[j] += (Fy - e->cy[j] * e->vy[j] - e->ky[j] * e->qy[j]) / e->my[j] * dt;
      e->qy[j] += e->vy[j] * dt; y += e->qy[j]; vys += e->vy[j];
    }
    if (truthy(wk)) {
      const double Fw = -(Fx * wnx + Fy * wny);
      wv += (Fw - wc * wv - wk * w) / wm * dt;
      w += wv * dt;
      if (w > 5e-4) { w = 5e-4; if (wv > 0) wv = 0; }
      else if (w < -2e-4) { w = -2e-4; if (wv < 0) wv = 0; }
      st[S_SUMW] += w; st[S_MAXW] = jmax(st[S_MAXW], fabs(w)); wSum += w;
    }
    if (++pr == spr) pr = 0;
    ph = pr % spt;
    if (ph == 0) {
      const double d0 = e->dR[(N - (pr / spt) % N) % N] * 1e-3;
      e->tsx[e->ts_i] = (float)(x - w * wnx); e->tsy[e->ts_i] = (float)(y - w * wny);
      e->ts_i = (e->ts_i + 1) % RN; e->ts_count++;
      e->wsx[e->ws_i] = (float)(x - w * wnx + d0 * vx); e->wsy[e->ws_i] = (float)(y - w * wny + d0 * vy);
      e->ws_i = (e->ws_i + 1) % RN; e->ws_count++;
    }
    if (ph == e->piStep) {
      const long long q = (long long)or_js_round((JS_PI - pr * dphi) / pitch);
      const int ip = (int)(((q % N) + N) % N);
      const double dp = e->dR[ip] * 1e-3;
      e->wpx[e->wsp_i] = (float)(x - w * wnx - dp * vx); e->wpy[e->wsp_i] = (float)(y - w * wny - dp * vy);
      e->wsp_i = (e->wsp_i + 1) % RN; e->wsp_count++;
    }
    t += dt;
    Fxy = or_hypot(Fx, Fy);
    e->runMaxF = jmax(e->runMaxF, Fxy);
    st[S_N]++; st[S_SUMF] += Fxy; st[S_MAXF] = jmax(st[S_MAXF], Fxy);
    st[S_SUMFX] += Fx; st[S_SUMFY] += Fy; st[S_SUMFZ] += Fz; st[S_MAXFZ] = jmax(st[S_MAXFZ], -Fz);
    st[S_SUMT] += T; st[S_MAXT] = jmax(st[S_MAXT], T);
    st[S_SUMX] += x; st[S_SUMY] += y;
    st[S_MINX] = jmin(st[S_MINX], x); st[S_MAXX] = jmax(st[S_MAXX], x);
    st[S_MINY] = jmin(st[S_MINY], y); st[S_MAXY] = jmax(st[S_MAXY], y);
    if (Fsum > 0) { st[S_SUMZC] += Fr2z / Fsum * Fxy; st[S_WZC] += Fxy; }
    if (++e->stepCount % e->decim == 0) {
      const int64_t jj = e->hist_i;
      e->hFx[jj] = (float)Fx; e->hFy[jj] = (float)Fy; e->hFz[jj] = (float)Fz;
      e->hx[jj] = (float)((x - w * wnx) * 1e6); e->hy[jj] = (float)((y - w * wny) * 1e6);
      e->hist_i = (jj + 1) % HN; e->hist_count++;
    }
  }
  e->x = e->x0 = x; e->y = e->y0 = y; e->vxs = e->vx0 = vxs; e->vys = e->vy0 = vys;
  e->t = t; e->pr = pr; e->w = w; e->wv = wv;
  e->runMeanW = nSteps ? wSum / nSteps : 0;
  return MC_OK;
}

void or_engine_stats(or_engine *e, mc_engine_stats *out, int take) {
  memcpy(out, e->st, sizeof(*out));
  if (take) new_stats(e->st);
}

int64_t or_engine_chip_events(const or_engine *e, mc_chip_event *out, int64_t cap) {
  int64_t i;
  for (i = 0; i < e->nev && i < cap; i++) out[i] = e->ev[i];
  return e->nev;
}

/* chatterIndex(), src/cutting.js:299-318 */
void or_engine_chatter_index(const or_engine *e, double out[3]) {
  const int64_t n = e->ts_count < 24 ? e->ts_count : 24;
  int64_t k, m;
  double mx = 0, my = 0, v = 0, ax = 0, ay = 0, w = 0, mm, r;
  if (n < 8) { out[0] = 0; out[1] = 0; out[2] = 0; return; }
  for (k = 0; k < n; k++) { const int64_t j = (e->ts_i - 1 - k + RN) % RN; mx += e->tsx[j]; my += e->tsy[j]; }
  mx /= n; my /= n;
  for (k = 0; k < n; k++) {
    const int64_t j = (e->ts_i - 1 - k + RN) % RN;
    const double dx = e->tsx[j] - mx, dy = e->tsy[j] - my;
    v += dx * dx + dy * dy;
  }
  out[2] = sqrt(v / n) * 1e6;
  r = or_js_round((double)n * e->spt / e->decim);
  m = (double)e->hist_count < r ? e->hist_count : (int64_t)r;
  for (k = 0; k < m; k++) { const int64_t j = (e->hist_i - 1 - k + HN) % HN; ax += e->hx[j]; ay += e->hy[j]; }
  mm = m ? (double)m : 1;
  ax /= mm; ay /= mm;
  for (k = 0; k < m; k++) {
    const int64_t j = (e->hist_i - 1 - k + HN) % HN;
    const double dx = e->hx[j] - ax, dy = e->hy[j] - ay;
    w += dx * dx + dy * dy;
  }
  out[1] = sqrt(w / mm);
  out[0] = out[2] / (out[1] + 0.05);
}

void or_engine_read_rings(const or_engine *e, float *hist, float *ts, float *ws, float *wsp) {
  if (hist) {
    memcpy(hist, e->hFx, sizeof(e->hFx)); memcpy(hist + HN, e->hFy, sizeof(e->hFy));
    memcpy(hist + 2 * HN, e->hFz, sizeof(e->hFz)); memcpy(hist + 3 * HN, e->hx, sizeof(e->hx));
    memcpy(hist + 4 * HN, e->hy, sizeof(e->hy));
  }
  if (ts) { memcpy(ts, e->tsx, sizeof(e->tsx)); memcpy(ts + RN, e->tsy, sizeof(e->tsy)); }
  if (ws) { memcpy(ws, e->wsx, sizeof(e->wsx)); memcpy(ws + RN, e->wsy, sizeof(e->wsy)); }
  if (wsp) { memcpy(wsp, e->wpx, sizeof(e->wpx)); memcpy(wsp + RN, e->wpy, sizeof(e->wpy)); }
}

double or_engine_dt(const or_engine *e) { return e->dt; }
int or_engine_flutes(const or_engine *e) { return e->N; }
int or_engine_spt(const or_engine *e) { return e->spt; }
double or_engine_run_max_f(const or_engine *e) { return e->runMaxF; }
double or_engine_dz(const or_engine *e) { return e->dz; }
int or_engine_max_slices(const or_engine *e) { return e->maxSlices; }
