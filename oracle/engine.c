/*
 * oracle/engine.c — TEST INFRASTRUCTURE ONLY (the oracle).  CPU restatement, in FP64
 * with the reference's float32 store points and summation order, of:
 *
 *   CuttingEngine.configure      src/cutting.js:39-82
 *   CuttingEngine.setEngagement  src/cutting.js:98-124   (flat and ball; no rAt profile)
 *   CuttingEngine.run            src/cutting.js:146-296  (side edges; no end edges, no wall)
 *   CuttingEngine.chatterIndex   src/cutting.js:299-318
 *   chatterAt / timeDomainLimit  src/stabtd.js:14-32
 *
 * Math.sin/cos/tan/pow come from fdlibm.c (V8's fdlibm port restated); JavaScript's
 * % on doubles is C fmod; (x)**2 is x*x (fdlibm pow special-cases y = 2); |0 on a
 * non-negative double is truncation; Math.round is or_js_round.  Float32Array stores
 * round to nearest float.  Compiled with -ffp-contract=off.
 *
 * Not restated (unused by chatterAt): end-edge forces (fza, src/cutting.js:215-237),
 * the 1-DOF wall (src/cutting.js:256-264), chip events, wall samples, torque/Fz stats.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define BINS 180
#define H_REF 0.05
#define H_MIN 0.002
#define JS_PI 3.141592653589793
#define HIST_N 4096
#define TS_N 64

typedef struct {
  int N, ball;
  double R, tanB, fluteLen;
  double dR[MC_MAX_FLUTES];
  double S, omega, Ttooth, dt;
  int spt, spr, decim;
  int nmx, nmy;
  double kx[MC_MAX_MODES], mx[MC_MAX_MODES], cx[MC_MAX_MODES], qx[MC_MAX_MODES], vx[MC_MAX_MODES];
  double ky[MC_MAX_MODES], my[MC_MAX_MODES], cy[MC_MAX_MODES], qy[MC_MAX_MODES], vy[MC_MAX_MODES];
  double dz;
  int maxSlices;
  float *surf;
  int pr;
  double x, y, vxs, vys, x0, y0, vx0, vy0;
  int64_t stepCount;
  int nonlinear;
  mc_material mat;
  /* engagement */
  double ux, uy, ex, ey, fz, tipZ, db;
  int nz;
  double sB[MC_MAX_SLICES], sZ[MC_MAX_SLICES], sLag[MC_MAX_SLICES], sR[MC_MAX_SLICES],
      sKap[MC_MAX_SLICES];
  float Hb[BINS];
  /* rings */
  float hx[HIST_N], hy[HIST_N];
  int hi;
  int64_t hcount;
  float tsx[TS_N], tsy[TS_N];
  int ti;
  int64_t tcount;
} Eng;

static double js_max(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == 0 && b == 0) return signbit(a) ? b : a;
  return a > b ? a : b;
}
static double js_min(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == 0 && b == 0) return signbit(a) ? a : b;
  return a < b ? a : b;
}

/* configure(tool, modal, mat, S, fnMax = 0), src/cutting.js:39-82 */
static int eng_configure(Eng *e, const mc_tool *tool, double fluteLen, double runout,
                         double runoutPhase, int nonlinear, const mc_cell *c, double S,
                         double fnMax) {
  int i;
  double fTop;
  memset(e, 0, sizeof(*e));
  e->N = tool->N;
  e->R = tool->R;
  e->ball = tool->type == MC_TOOL_BALL;
  e->fluteLen = fluteLen;
  e->tanB = or_tan(tool->helixDeg * JS_PI / 180);
  for (i = 0; i < e->N; i++) e->dR[i] = runout * or_cos(2 * JS_PI * i / e->N - runoutPhase);
  e->S = js_max(100, S);
  e->omega = e->S * 2 * JS_PI / 60;
  e->Ttooth = 60 / (e->S * e->N);
  fTop = -INFINITY;
  for (i = 0; i < c->nmx; i++) fTop = js_max(fTop, c->mx[i].fn);
  for (i = 0; i < c->nmy; i++) fTop = js_max(fTop, c->my[i].fn);
  e->spt = (int)js_max(48, ceil(e->Ttooth * js_max(fTop, fnMax) * 24));
  e->dt = e->Ttooth / e->spt;
  e->nmx = c->nmx;
  e->nmy = c->nmy;
  for (i = 0; i < c->nmx; i++) {
    double w = 2 * JS_PI * c->mx[i].fn;
    e->kx[i] = c->mx[i].k * 1e3;
    e->mx[i] = e->kx[i] / (w * w);
    e->cx[i] = 2 * c->mx[i].zeta * sqrt(e->kx[i] * e->mx[i]);
  }
  for (i = 0; i < c->nmy; i++) {
    double w = 2 * JS_PI * c->my[i].fn;
    e->ky[i] = c->my[i].k * 1e3;
    e->my[i] = e->ky[i] / (w * w);
    e->cy[i] = 2 * c->my[i].zeta * sqrt(e->ky[i] * e->my[i]);
  }
  e->dz = js_max(0.25, fluteLen / 48);
  e->maxSlices = (int)ceil(fluteLen / e->dz);
  e->spr = e->N * e->spt;
  e->surf = (float *)calloc((size_t)e->maxSlices * e->spr, sizeof(float));
  if (!e->surf) return -1;
  e->decim = (int)js_max(1, or_js_round(e->spt / 40.0));
  e->nonlinear = nonlinear;
  e->mat = c->mat;
  return 0;
}

/* setEngagement(u, fz, tipZ, ap), src/cutting.js:98-124 (flat and ball) */
static void eng_set_engagement(Eng *e, double u0, double u1, double fz, double tipZ, double ap) {
  int k;
  e->ux = u0; e->uy = u1; e->ex = -u1; e->ey = u0;
  e->fz = fz; e->tipZ = tipZ;
  e->nz = ap <= 0 ? 0 : (int)js_min(e->maxSlices, ceil(ap / e->dz));
  e->db = e->dz;
  for (k = 0; k < e->nz; k++) {
    double zr, r = e->R, kap = 1;
    e->sB[k] = js_min(e->dz, ap - k * e->dz);
    zr = k * e->dz + 0.5 * e->sB[k];
    e->sZ[k] = tipZ + zr;
    if (e->ball && zr < e->R) {
      double t = e->R - zr;
      r = sqrt(js_max(1e-6, e->R * e->R - t * t));
      kap = r / e->R;
    }
    e->sR[k] = r; e->sKap[k] = kap;
    e->sLag[k] = zr * e->tanB / e->R;
  }
}

/* run(nSteps), src/cutting.js:146-296, side edges only (fza = 0, rigid workpiece).
 * Optional per-step trace of Fx, Fy, x, y. */
static void eng_run(Eng *e, int64_t nSteps, or_counts *cnt, double *tFx, double *tFy,
                    double *tx, double *ty) {
  const int N = e->N, spt = e->spt, spr = e->spr, nz = e->nz;
  const double ux = e->ux, uy = e->uy, vx = e->ex, vy = e->ey, fz = e->fz, dt = e->dt;
  const double Ktc = e->mat.Ktc, Krc = e->mat.Krc, Kte = e->mat.Kte, Kre = e->mat.Kre;
  const int nl = e->nonlinear;
  const double mc = nl ? e->mat.mc : 0;
  const double pdK = nl ? e->mat.Ci * e->db * 1e-3 / (e->omega * 1e-3) : 0;
  const double TWO_PI = 2 * JS_PI, binScale = BINS / TWO_PI, pitch = TWO_PI / N;
  const double dphi = e->omega * dt;
  const double hMax = js_max(0.1, 3 * fz);
  const double w = 0, wv = 0, wnx = 1, wny = 0; /* setWork(null) */
  double x, y, vxs, vys;
  int pr, i, j, k;
  int64_t s;
  e->qx[0] += e->x - e->x0; e->qy[0] += e->y - e->y0;
  e->vx[0] += e->vxs - e->vx0; e->vy[0] += e->vys - e->vy0;
  x = e->x; y = e->y; vxs = e->vxs; vys = e->vys; pr = e->pr;
  for (s = 0; s < nSteps; s++) {
    const double theta = pr * dphi;
    const double rx = x - w * wnx, ry = y - w * wny;
    const double xm = rx * 1e3, ym = ry * 1e3;
    double Fx = 0, Fy = 0, Fxy;
    const double rvx = vxs - wv * wnx, rvy = vys - wv * wny;
    for (i = 0; i < N; i++) {
      const double dRi = e->dR[i];
      int pi = pr + i * spt;
      if (pi >= spr) pi -= spr;
      for (k = 0; k < nz; k++) {
        double phi = fmod(theta + i * pitch - e->sLag[k], TWO_PI);
        double sn, cs, nxg, nyg, rn, prev, feedH, h, se, bk, Ft, Fr, txg, tyg, hb;
        int idx, bin;
        if (phi < 0) phi += TWO_PI;
        sn = or_sin(phi); cs = or_cos(phi);
        nxg = sn * ux + cs * vx; nyg = sn * uy + cs * vy;
        rn = xm * nxg + ym * nyg + dRi;
        idx = k * spr + pi;
        prev = e->surf[idx];
        feedH = fz * sn * e->sKap[k];
        bin = (int)(phi * binScale);
        hb = (bin >= 0 && bin < BINS) ? (double)e->Hb[bin] : NAN; /* undefined <= z is false */
        if (cnt) cnt->visits++;
        if (hb <= e->sZ[k] || sn <= 0) { e->surf[idx] = (float)rn; continue; }
        h = feedH + rn - prev;
        if (h <= 0) { e->surf[idx] = (float)(prev - feedH); continue; }
        e->surf[idx] = (float)rn;
        if (cnt) cnt->engaged++;
        if (h > hMax) h = hMax;
        se = mc != 0 ? or_pow(H_REF / (h > H_MIN ? h : H_MIN), mc) : 1;
        bk = e->sB[k];
        Ft = Ktc * se * h * bk + Kte * bk;
        Fr = Krc * se * h * bk + Kre * bk;
        if (pdK != 0 && !isnan(pdK)) Fr += pdK * (bk / e->db) / e->sR[k] * (rvx * nxg + rvy * nyg);
        txg = cs * ux - sn * vx; tyg = cs * uy - sn * vy;
        Fx += -Ft * txg - Fr * nxg;
        Fy += -Ft * tyg - Fr * nyg;
      }
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
    if (++pr == spr) pr = 0;
    if (pr % spt == 0) {
      e->tsx[e->ti] = (float)(x - w * wnx); e->tsy[e->ti] = (float)(y - w * wny);
      e->ti = (e->ti + 1) % TS_N; e->tcount++;
    }
    Fxy = or_hypot(Fx, Fy);
    if (cnt) {
      cnt->steps++;
      cnt->sumF += Fxy;
      cnt->maxF = js_max(cnt->maxF, Fxy);
    }
    if (++e->stepCount % e->decim == 0) {
      int jj = e->hi;
      e->hx[jj] = (float)((x - w * wnx) * 1e6); e->hy[jj] = (float)((y - w * wny) * 1e6);
      e->hi = (jj + 1) % HIST_N; e->hcount++;
    }
    if (tFx) tFx[s] = Fx;
    if (tFy) tFy[s] = Fy;
    if (tx) tx[s] = x;
    if (ty) ty[s] = y;
  }
  e->x = e->x0 = x; e->y = e->y0 = y; e->vxs = e->vx0 = vxs; e->vys = e->vy0 = vys; e->pr = pr;
}

/* chatterIndex(), src/cutting.js:299-318 */
static void eng_chatter_index(const Eng *e, double *index, double *amp, double *once) {
  int n = (int)(e->tcount < 24 ? e->tcount : 24), k, j;
  int64_t m;
  double mx = 0, my = 0, v = 0, sOnce, ax = 0, ay = 0, w = 0, sAll, mm;
  if (n < 8) { *index = 0; *amp = 0; *once = 0; return; }
  for (k = 0; k < n; k++) { j = (e->ti - 1 - k + TS_N) % TS_N; mx += e->tsx[j]; my += e->tsy[j]; }
  mx /= n; my /= n;
  for (k = 0; k < n; k++) {
    double dx, dy;
    j = (e->ti - 1 - k + TS_N) % TS_N;
    dx = e->tsx[j] - mx; dy = e->tsy[j] - my;
    v += dx * dx + dy * dy;
  }
  sOnce = sqrt(v / n) * 1e6;
  {
    double r = or_js_round((double)n * e->spt / e->decim);
    m = e->hcount < (int64_t)r ? e->hcount : (int64_t)r;
  }
  for (k = 0; k < m; k++) { j = (int)((e->hi - 1 - k + HIST_N) % HIST_N); ax += e->hx[j]; ay += e->hy[j]; }
  mm = m ? (double)m : 1;
  ax /= mm; ay /= mm;
  for (k = 0; k < m; k++) {
    double dx, dy;
    j = (int)((e->hi - 1 - k + HIST_N) % HIST_N);
    dx = e->hx[j] - ax; dy = e->hy[j] - ay;
    w += dx * dx + dy * dy;
  }
  sAll = sqrt(w / mm);
  *index = sOnce / (sAll + 0.05);
  *amp = sAll;
  *once = sOnce;
}

/* the engine chatterAt builds: src/stabtd.js:15-19 */
static int trial_setup(Eng *e, const mc_cell *c, double ap) {
  int b;
  if (eng_configure(e, &c->tool, js_max(c->tool.fluteLen, ap + 1), 0, 0, c->nonlinear, c, c->S, 0))
    return -1;
  for (b = 0; b < BINS; b++) {
    double phi = (b + 0.5) / BINS * 2 * JS_PI;
    e->Hb[b] = (phi >= c->phiSt && phi <= c->phiEx) ? 0.0f : -1e3f;
  }
  eng_set_engagement(e, 1, 0, c->fz, -ap, ap);
  e->vxs = 0.01;
  return 0;
}

static int cell_ok(const mc_cell *c) {
  return c && c->tool.N >= 1 && c->tool.N <= MC_MAX_FLUTES && c->nmx >= 1 &&
         c->nmx <= MC_MAX_MODES && c->nmy >= 1 && c->nmy <= MC_MAX_MODES;
}

int or_chatter_at(const mc_cell *c, double ap, mc_trial *out, or_counts *cnt) {
  Eng *e;
  double idx, amp, once;
  int64_t n;
  if (!cell_ok(c)) return MC_ERR_ARG;
  e = (Eng *)malloc(sizeof(Eng));
  if (!e) return MC_ERR_NOMEM;
  if (trial_setup(e, c, ap)) { free(e); return MC_ERR_NOMEM; }
  if (e->nz > MC_MAX_SLICES) { free(e->surf); free(e); return MC_ERR_LIMIT; }
  n = (int64_t)e->spt * c->tool.N * c->revs;
  eng_run(e, n, cnt, NULL, NULL, NULL, NULL);
  eng_chatter_index(e, &idx, &amp, &once);
  if (out) {
    out->index = idx; out->amp = amp; out->once = once;
    out->chatter = idx > MC_CHATTER_IDX;
    out->nz = e->nz;
    out->steps = n;
  }
  free(e->surf);
  free(e);
  return MC_OK;
}

static int chatter(const mc_cell *c, double ap, mc_limit *acc, int *st) {
  mc_trial t;
  or_counts cnt;
  memset(&cnt, 0, sizeof(cnt));
  *st = or_chatter_at(c, ap, &t, &cnt);
  acc->trials++;
  acc->steps += cnt.steps;
  acc->visits += cnt.visits;
  acc->engaged += cnt.engaged;
  return t.chatter;
}

/* timeDomainLimit, src/stabtd.js:24-32 */
int or_time_domain_limit(const mc_cell *c, const mc_bisect_opts *o, mc_limit *out) {
  double lo, hi;
  int i, st = 0;
  memset(out, 0, sizeof(*out));
  if (!chatter(c, o->aMax, out, &st)) {
    out->alim = o->aMax; out->lo = o->aMax; out->hi = INFINITY; out->open = 1;
    return st;
  }
  lo = o->aMin; hi = o->aMax;
  for (i = 0; i < o->iters && st == MC_OK; i++) {
    double m = sqrt(lo * hi);
    if (chatter(c, m, out, &st)) hi = m; else lo = m;
  }
  out->alim = sqrt(lo * hi); out->lo = lo; out->hi = hi; out->open = 0;
  return st;
}

/* bracketed search from a hint (no reference counterpart): check both ends, bisect
 * geometrically until hi - lo <= tol (at most iters halvings); a failed hint falls back
 * to timeDomainLimit, counting the trials already spent */
int or_bracket_limit(const mc_cell *c, const mc_bisect_opts *o, double lo_hint, double hi_hint,
                     double tol, mc_limit *out, int32_t *fell_back) {
  double lo = lo_hint, hi = hi_hint;
  int i, st = 0;
  mc_limit spent;
  memset(out, 0, sizeof(*out));
  if (fell_back) *fell_back = 0;
  if (!(lo_hint > 0) || !(hi_hint > lo_hint) || hi_hint > o->aMax || lo_hint < o->aMin) goto fallback;
  if (chatter(c, lo, out, &st)) goto fallback;
  if (!chatter(c, hi, out, &st)) goto fallback;
  for (i = 0; i < o->iters && hi - lo > tol && st == MC_OK; i++) {
    double m = sqrt(lo * hi);
    if (chatter(c, m, out, &st)) hi = m; else lo = m;
  }
  out->alim = sqrt(lo * hi); out->lo = lo; out->hi = hi; out->open = 0;
  return st;
fallback:
  spent = *out;
  if (fell_back) *fell_back = 1;
  st = or_time_domain_limit(c, o, out);
  out->trials += spent.trials; out->steps += spent.steps; out->visits += spent.visits;
  out->engaged += spent.engaged;
  return st;
}

int64_t or_engine_trace(const mc_cell *c, double ap, int64_t nsteps, double *Fx, double *Fy,
                        double *x, double *y) {
  Eng *e;
  int64_t total;
  if (!cell_ok(c)) return -1;
  e = (Eng *)malloc(sizeof(Eng));
  if (!e) return -1;
  if (trial_setup(e, c, ap)) { free(e); return -1; }
  total = (int64_t)e->spt * c->tool.N * c->revs;
  if (nsteps > total) nsteps = total;
  eng_run(e, nsteps, NULL, Fx, Fy, x, y);
  free(e->surf);
  free(e);
  return nsteps;
}

/* the engine of one simulator sub-step: configure with the tool's own flute length and
 * runout (src/sim.js:252-257), setEngagement(u, fz, tipZ, ap) with the caller's map
 * (src/sim.js:403-404), a velocity set from outside, then run step by step */
int64_t or_engine_trace_ex(const mc_cell *c, const mc_engage_spec *s, int64_t nsteps, double *Fx,
                           double *Fy, double *x, double *y) {
  Eng *e;
  if (!cell_ok(c) || !s || nsteps < 0) return -1;
  e = (Eng *)malloc(sizeof(Eng));
  if (!e) return -1;
  if (eng_configure(e, &c->tool, c->tool.fluteLen, c->tool.runout, c->tool.runoutPhase,
                    c->nonlinear, c, c->S, s->fnMax)) {
    free(e);
    return -1;
  }
  memcpy(e->Hb, s->Hb, sizeof(e->Hb));
  eng_set_engagement(e, s->ux, s->uy, s->fz, s->tipZ, s->ap);
  if (e->nz > MC_MAX_SLICES) { free(e->surf); free(e); return -1; }
  e->vxs = s->vx0;
  e->vys = s->vy0;
  eng_run(e, nsteps, NULL, Fx, Fy, x, y);
  free(e->surf);
  free(e);
  return nsteps;
}

/* ---- host-thread batches (CPU baseline): one cell per thread at a time ---- */
typedef struct {
  const mc_cell *cells;
  const double *ap;
  const mc_bisect_opts *o;
  mc_limit *lim;
  mc_trial *tri;
  int n;
  int next;
  pthread_mutex_t mu;
} Job;

static void *worker(void *arg) {
  Job *jb = (Job *)arg;
  for (;;) {
    int i;
    pthread_mutex_lock(&jb->mu);
    i = jb->next++;
    pthread_mutex_unlock(&jb->mu);
    if (i >= jb->n) break;
    if (jb->lim) or_time_domain_limit(&jb->cells[i], jb->o, &jb->lim[i]);
    else or_chatter_at(&jb->cells[i], jb->ap[i], &jb->tri[i], NULL);
  }
  return NULL;
}

static int run_jobs(Job *jb, int nthreads) {
  pthread_t th[256];
  int t;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_mutex_init(&jb->mu, NULL);
  for (t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker, jb);
  for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&jb->mu);
  return MC_OK;
}

int or_bisect_batch_mt(const mc_cell *cells, int n, const mc_bisect_opts *o, int nthreads,
                       mc_limit *out) {
  Job jb;
  memset(&jb, 0, sizeof(jb));
  jb.cells = cells; jb.o = o; jb.lim = out; jb.n = n;
  return run_jobs(&jb, nthreads);
}

int or_chatter_batch_mt(const mc_cell *cells, const double *ap, int n, int nthreads,
                        mc_trial *out) {
  Job jb;
  memset(&jb, 0, sizeof(jb));
  jb.cells = cells; jb.ap = ap; jb.tri = out; jb.n = n;
  return run_jobs(&jb, nthreads);
}

void or_math_batch(int fn, const double *x, const double *y, double *out, int64_t n) {
  int64_t i;
  for (i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = or_sin(x[i]); break;
      case 1: out[i] = or_cos(x[i]); break;
      case 2: out[i] = or_tan(x[i]); break;
      case 3: out[i] = or_pow(x[i], y[i]); break;
      case 4: out[i] = or_hypot(x[i], y[i]); break;
      case 5: out[i] = or_atan(x[i]); break;
      case 6: out[i] = or_atan2(x[i], y[i]); break;
      default: out[i] = NAN;
    }
  }
}

/* toolModel for a uniform beam, src/tool.js:16-32 */
void or_tool_model(double D, double stickout, double kSpindle, double mHolder, double *k,
                   double *fn) {
  const double E = 600000, RHO = 14500;
  double L = stickout, De = 0.8 * D;
  double I = JS_PI * or_pow(De, 4) / 64;
  double kTool = 3 * E * I / or_pow(L, 3);
  double kTotal = 1 / (1 / kTool + 1 / kSpindle);
  double EI = E * 1e6 * I * 1e-12;
  double t = De * 1e-3, t2 = L * 1e-3;
  double rhoA = RHO * JS_PI * (t * t) / 4;
  double wTool = (1.875 * 1.875) * sqrt(EI / rhoA) / (t2 * t2);
  double mEff = kTool * 1e3 / (wTool * wTool) + mHolder;
  *k = kTotal;
  *fn = sqrt(kTotal * 1e3 / mEff) / (2 * JS_PI);
}
