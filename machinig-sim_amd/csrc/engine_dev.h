// engine_dev.h — the device-resident CuttingEngine record and its step loop
// (src/cutting.js:21-326), shared by the engine's own entry points (engine.hip) and the
// simulator's sub-step pipeline (simrun.hip).  Library-internal; not part of the ABI.
//
// Per step, the lanes of one wavefront split the (tooth, slice) side-edge visits and the
// (tooth, ring) end-edge visits; each lane evaluates its visits' engagement, chip thickness
// and forces, writes the delay line and leaves every quantity the reference accumulates in
// LDS, at the visit's position in the reference's loop order (tooth by tooth, its slices,
// then its end rings).  Nine lanes then each sum one of the accumulations (Fx, Fy, Fz,
// torque, force sum, its moment, and the three engagement statistics) serially in that
// order, and one lane per tooth sums the tooth's chip volume and decides its chip event, so
// every FP64 sum is the reference's, bit for bit.  The modal and wall updates run
// redundantly on all lanes; lane 0 keeps the step statistics and writes the sample rings.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mcsim.h"
#include "mc_math.h"

namespace mce {

constexpr int EBINS = MC_END_BINS, ERINGS = MC_END_RINGS;
constexpr int HN = MC_ENGINE_HIST, RN = MC_ENGINE_RING;
constexpr int MS = MC_MAX_SLICES;
constexpr int SLOT_CAP = MC_MAX_FLUTES * (MS + ERINGS);
constexpr double JS_PI = 3.141592653589793;
constexpr double H_REF = 0.05, H_MIN = 0.002;

enum { S_N, S_SUMF, S_MAXF, S_SUMFX, S_SUMFY, S_SUMFZ, S_MAXFZ, S_SUMT, S_MAXT, S_SUMX, S_SUMY, S_MINX,
       S_MAXX, S_MINY, S_MAXY, S_SUMZC, S_WZC, S_SUMW, S_MAXW, S_SUMHB, S_SUMB, S_SUMRB, S_ENGT, S_COUNT };

// accumulations kept per visit: the step's sums (lanes 0..T_SUMS-1 of the chain; the last
// three carry across steps), then the tooth's chip volume, angle and height
enum { T_FX, T_FY, T_FZ, T_TQ, T_HB, T_B, T_RB, T_VOL, T_PHI, T_Z, T_COUNT };
constexpr int T_SUMS = T_RB + 1;
// per-step records of a chunk, for the statistics taken after it
// (R_NFZ holds -Fz and R_AW |w|, the operands of the Fz and wall maxima; R_ONE is 1.0 throughout,
// the step count's operand and the condition of the unconditional fields)
enum { R_FXY, R_FX, R_FY, R_FZ, R_TQ, R_X, R_Y, R_W, R_ENG, R_ZC, R_ZCOK, R_NFZ, R_AW, R_ONE, R_COUNT };

// One engine.  The host keeps a mirror, uploads it before a run and reads it back after.
struct EngDev {
  // configure
  double R, tanB, omega, dt, Ttooth, dz, S;
  double kx[MC_MAX_MODES], mx[MC_MAX_MODES], cx[MC_MAX_MODES], qx[MC_MAX_MODES], vxm[MC_MAX_MODES];
  double ky[MC_MAX_MODES], my[MC_MAX_MODES], cy[MC_MAX_MODES], qy[MC_MAX_MODES], vym[MC_MAX_MODES];
  double dR[MC_MAX_FLUTES];
  int32_t N, spt, spr, decim, nmx, nmy, maxSlices, piStep;
  int32_t pr, nz, pad0, pad1;
  int64_t stepCount;
  // run constants, derived on the host from the material, wear and engagement
  double Ktc, Krc, Kac, Kae, Kte, Kre, mc, pdK, dphi, hMax, pitch, binScale;
  // state
  double x, y, vx, vy, x0, y0, vx0, vy0, t, w, wv, runMeanW, wk, wm, wc, wnx, wny;
  // setEngagement / setEndEngagement
  double ux, uy, ex, ey, fz, tipZ, db;
  double sB[MS], sZ[MS], sLag[MS], sR[MS], sKap[MS];
  double fza, endR, ringR[ERINGS], ringCos[ERINGS], ringZ[ERINGS], ringDs[ERINGS];
  float Hb[MC_BINS], Eb[EBINS * ERINGS];
  // per tooth
  double toothVol[MC_MAX_FLUTES], toothPhi[MC_MAX_FLUTES], toothZ[MC_MAX_FLUTES];
  int32_t toothEngaged[MC_MAX_FLUTES];
  // stats and rings
  double st[S_COUNT];
  int64_t hist_i, hist_count, ts_i, ts_count, ws_i, ws_count, wsp_i, wsp_count;
  // this run
  int64_t nsteps, nev, ev_cap;
  double runMaxF;  // max |(Fx, Fy)| over the last run
  float* surf;   // [spr][maxSlices]
  float* hist;   // [5][HN]: Fx Fy Fz x y
  float* rings;  // [6][RN]: ts x, ts y, ws x, ws y, wsp x, wsp y
  double* ev;    // [ev_cap][3]: vol, phi, z
};

// the step loop runs on one wavefront: its lanes' LDS and delay-line accesses are ordered at
// wavefront scope (a workgroup-scope fence would wait for every store to leave the CU)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Math.max / Math.min: NaN wins, +0 is above -0
__host__ __device__ __forceinline__ double jmax(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a == 0 && b == 0) return __builtin_signbit(a) ? b : a;
  return a > b ? a : b;
}
__host__ __device__ __forceinline__ double jmin(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a == 0 && b == 0) return __builtin_signbit(a) ? a : b;
  return a < b ? a : b;
}
__host__ __device__ __forceinline__ bool truthy(double v) { return v != 0 && v == v; }
// the same as selects (no branches): for the statistics' per-lane operations
__device__ __forceinline__ double jmax_sel(double a, double b) {
  double r = a > b ? a : b;
  r = (a == 0 && b == 0) ? (__builtin_signbit(a) ? b : a) : r;
  return (a != a || b != b) ? __builtin_nan("") : r;
}
__device__ __forceinline__ double jmin_sel(double a, double b) {
  double r = a < b ? a : b;
  r = (a == 0 && b == 0) ? (__builtin_signbit(a) ? a : b) : r;
  return (a != a || b != b) ? __builtin_nan("") : r;
}

// a chunk of steps is staged in LDS: the edge angle with its sine and cosine of every
// (step, visit), and the delay-line cells of its slice visits; GCAP entries at most
constexpr int GCAP = 1024;

// LDS of one engine's wavefront: the per-visit accumulations and the step sums
struct EngLds {
  double term[T_COUNT][SLOT_CAP];
  int teng[SLOT_CAP];
  double res[T_SUMS];
  // the run's constants, staged from the engine record: the slice tables, the edge force
  // terms Kte b, Kre b (slices, then rings with ds), Kae b, the process-damping factor
  // pdK (b / db) / r of each slice
  double sLag[MS], sKap[MS], sZ[MS], sB[MS], sR[MS];
  double cTe[MS + ERINGS], cRe[MS + ERINGS], cAe[MS], cPd[MS];
  double dR[MC_MAX_FLUTES];
  double ring[4 * ERINGS];          // ringR ringCos ringZ ringDs
  float Hb[MC_BINS], Eb[EBINS * ERINGS];
  double gphi[GCAP], gsn[GCAP], gcs[GCAP];
  float win[GCAP];
  // a chunk's tangential and radial force of every (step, visit), for the force magnitudes
  // taken after it, and its per-step records
  double dFt[GCAP], dFr[GCAP];
  double rec[R_COUNT][64];
};

// the rare paths of the step loop, kept out of line so that the loop's code stays small:
// pow_js where the branch-free pow_cell does not apply, hypot_js of non-finite operands
__device__ __attribute__((noinline)) double pow_js_outlined(double x, double y) { return mcm::pow_js(x, y); }
__device__ __attribute__((noinline)) double hypot_js_outlined(double a, double b) { return mcm::hypot_js(a, b); }

// Math.hypot(a, b) for finite a, b without branches: V8's scaled, compensated sum (as
// mcm::hypot_js), which for two terms is sqrt((|a|/m)^2 + (|b|/m)^2) m with m = max(|a|, |b|)
// (the compensation term is exactly 0 after the first); non-finite operands go to hypot_js
__device__ __forceinline__ double hypot_finite(double a, double b) {
  const double fa = __builtin_fabs(a), fb = __builtin_fabs(b);
  if (!(fa < __builtin_inf() && fb < __builtin_inf())) return hypot_js_outlined(a, b);
  const double mx = fa > fb ? fa : fb;
  const double m = mx == 0 ? 1.0 : mx;
  const double na = fa / m, nb = fb / m;
  const double r = __builtin_sqrt(na * na + nb * nb) * m;
  return mx == 0 ? 0.0 : r;
}

// a double of lane `src`, broadcast to every lane
__device__ __forceinline__ double read_lane(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// chatterIndex() (src/cutting.js:299-318) over the engine's rings, by one thread: the RMS
// scatter of the last n <= 24 once-per-tooth samples about their mean, over the RMS of the
// last min(count, round(n spt / decim)) history samples about theirs, both newest first
__device__ __forceinline__ void eng_chatter_index(const EngDev& E, double* out) {
  const int64_t n = E.ts_count < 24 ? E.ts_count : 24;
  if (n < 8) { out[0] = 0; out[1] = 0; out[2] = 0; return; }
  const float* tsx = E.rings;
  const float* tsy = E.rings + RN;
  double mx = 0, my = 0, v = 0;
  for (int64_t k = 0; k < n; k++) { const int64_t j = (E.ts_i - 1 - k + RN) % RN; mx += tsx[j]; my += tsy[j]; }
  mx /= n; my /= n;
  for (int64_t k = 0; k < n; k++) {
    const int64_t j = (E.ts_i - 1 - k + RN) % RN;
    const double dx = tsx[j] - mx, dy = tsy[j] - my;
    v += dx * dx + dy * dy;
  }
  const double once = __builtin_sqrt(v / n) * 1e6;
  const double r = mcm::round_js((double)n * E.spt / E.decim);
  const int64_t m = (double)E.hist_count < r ? E.hist_count : (int64_t)r;
  const float* hx = E.hist + 3 * HN;
  const float* hy = E.hist + 4 * HN;
  double ax = 0, ay = 0, w = 0;
  for (int64_t k = 0; k < m; k++) { const int64_t j = (E.hist_i - 1 - k + HN) % HN; ax += hx[j]; ay += hy[j]; }
  const double mm = m ? (double)m : 1.0;
  ax /= mm; ay /= mm;
  for (int64_t k = 0; k < m; k++) {
    const int64_t j = (E.hist_i - 1 - k + HN) % HN;
    const double dx = hx[j] - ax, dy = hy[j] - ay;
    w += dx * dx + dy * dy;
  }
  const double amp = __builtin_sqrt(w / mm);
  out[0] = once / (amp + 0.05);
  out[1] = amp;
  out[2] = once;
}

// newStats() (src/cutting.js:323-326)
__host__ __device__ __forceinline__ void eng_new_stats(double* st) {
  for (int k = 0; k < S_COUNT; k++) st[k] = 0;
  st[S_MINX] = __builtin_inf(); st[S_MAXX] = -__builtin_inf();
  st[S_MINY] = __builtin_inf(); st[S_MAXY] = -__builtin_inf();
}

// run(nSteps) of src/cutting.js:146-296 on engine E by the calling wavefront.
//
// The steps run in chunks of at most one tooth period.  Before a chunk, every lane computes
// its share of the chunk's edge angles (the slice angle (theta + i pitch - lag) mod 2 pi and
// the end-edge angle, with their fdlibm sine and cosine) and loads the delay-line cells the
// chunk's slice visits read: within one tooth period no cell is visited twice, so the chunk
// reads and writes them in LDS and stores them back after it.  A step's own chain is then
// the chip thickness, the force law, the serial sums, the chip events and the modes (one
// lane per mode, summed in mode order).  What no later step reads is taken after the chunk,
// from the chunk's records, in the reference's order: the force magnitudes |(Ft, Fr)| and
// their sums, |(Fx, Fy)|, the statistics (one lane per field), the rings and the history.
__device__ __forceinline__ void engine_steps(EngDev& E, int64_t nSteps, EngLds& L) {
  auto& term = L.term;
  auto& teng = L.teng;
  auto& res = L.res;
  auto& rec = L.rec;
  const int lane = threadIdx.x & 63;
  const double TWO_PI = mcm::TWO_PI;

  const int N = E.N, spt = E.spt, spr = E.spr, nz = E.nz, MSL = E.maxSlices, decim = E.decim;
  const int nmx = E.nmx, nmy = E.nmy, piStep = E.piStep;
  const double ux = E.ux, uy = E.uy, vx = E.ex, vy = E.ey, fz = E.fz, dt = E.dt, fza = E.fza, tipZ = E.tipZ;
  const double Ktc = E.Ktc, Krc = E.Krc, Kac = E.Kac, Kae = E.Kae, Kte = E.Kte, Kre = E.Kre;
  const double mc = E.mc, pdK = E.pdK, db = E.db, dphi = E.dphi, hMax = E.hMax, pitch = E.pitch,
               binScale = E.binScale;
  const double wnx = E.wnx, wny = E.wny, wk = E.wk, wm = E.wm, wc = E.wc;
  const bool ends = truthy(fza);
  const int per = nz + (ends ? ERINGS : 0);
  const int NS = N * per;
  const bool mcOn = truthy(mc), pdOn = truthy(pdK), wallOn = truthy(wk);
  // the size effect's pow: the branch-free form where it equals pow_js (x = H_REF / h lies in
  // [H_REF / hMax, 25] and the exponent in pow_cell's domain), else pow_js itself
  const bool cellPow = mcOn && mcm::pow_cell_ok(mc) && hMax <= 1e27;
  const double my1 = mcm::clear_lo(mc), myr = mc - my1;
  double runMaxF = 0;
  float* const surf = E.surf;

  for (int k = lane; k < nz; k += 64) {
    const double bk = E.sB[k], sr = E.sR[k];
    L.sLag[k] = E.sLag[k]; L.sKap[k] = E.sKap[k]; L.sZ[k] = E.sZ[k]; L.sB[k] = bk; L.sR[k] = sr;
    L.cTe[k] = Kte * bk; L.cRe[k] = Kre * bk; L.cAe[k] = Kae * bk;
    L.cPd[k] = pdK * (bk / db) / sr;
  }
  for (int b = lane; b < MC_BINS; b += 64) L.Hb[b] = E.Hb[b];
  for (int b = lane; b < EBINS * ERINGS; b += 64) L.Eb[b] = E.Eb[b];
  if (lane < MC_MAX_FLUTES) L.dR[lane] = E.dR[lane];
  if (lane < ERINGS) {
    L.ring[lane] = E.ringR[lane]; L.ring[ERINGS + lane] = E.ringCos[lane];
    L.ring[2 * ERINGS + lane] = E.ringZ[lane]; L.ring[3 * ERINGS + lane] = E.ringDs[lane];
    L.cTe[MS + lane] = Kte * E.ringDs[lane]; L.cRe[MS + lane] = Kre * E.ringDs[lane];
  }

  // modes: lane j < nmx integrates x-mode j, lane MC_MAX_MODES + j y-mode j; a state set from
  // outside since the last run enters the first mode (src/cutting.js:164)
  const bool xmode = lane < nmx, ymode = lane >= MC_MAX_MODES && lane < MC_MAX_MODES + nmy;
  const int mj = lane < MC_MAX_MODES ? lane : (lane - MC_MAX_MODES) & (MC_MAX_MODES - 1);
  double mk = 0, mmass = 1, mcd = 0, mq = 0, mv = 0;
  if (xmode) { mk = E.kx[mj]; mmass = E.mx[mj]; mcd = E.cx[mj]; mq = E.qx[mj]; mv = E.vxm[mj]; }
  if (ymode) { mk = E.ky[mj]; mmass = E.my[mj]; mcd = E.cy[mj]; mq = E.qy[mj]; mv = E.vym[mj]; }
  if (lane == 0) { mq += E.x - E.x0; mv += E.vx - E.vx0; }
  if (lane == MC_MAX_MODES) { mq += E.y - E.y0; mv += E.vy - E.vy0; }
  double x = E.x, y = E.y, vxs = E.vx, vys = E.vy, t = E.t, w = E.w, wv = E.wv, wSum = 0;
  int pr = E.pr;
  int ph = pr % spt, tq = pr / spt;   // pr % stepsPerTooth and pr / stepsPerTooth, kept as counters
  int64_t stepCount = E.stepCount;
  int dleft = decim - (int)(stepCount % decim);   // steps until ++stepCount % decim == 0
  // sum lanes T_HB..T_RB carry the engagement statistics across steps; tooth lanes their tooth
  const int tooth = lane - 16;
  const bool isTooth = tooth >= 0 && tooth < N;
  double chain = lane == T_HB ? E.st[S_SUMHB] : lane == T_B ? E.st[S_SUMB] : lane == T_RB ? E.st[S_SUMRB] : 0.0;
  double tvol = isTooth ? E.toothVol[tooth] : 0.0, tphi = isTooth ? E.toothPhi[tooth] : 0.0,
         tz = isTooth ? E.toothZ[tooth] : 0.0;
  int tEng = isTooth ? E.toothEngaged[tooth] : 0;
  // the step statistics: lane f keeps field f (the three engagement sums are the sum lanes');
  // each lane adds, maximises or minimises its own operand of each step, a record field
  enum { OP_NONE, OP_ADD, OP_MAX, OP_MIN };
  int op = OP_NONE, src = 0, cond = 0;   // cond 1: Fsum > 0, 2: the wall mode is on
  switch (lane) {
    case S_N: op = OP_ADD; src = -1; break;
    case S_SUMF: op = OP_ADD; src = R_FXY; break;
    case S_MAXF: op = OP_MAX; src = R_FXY; break;
    case S_SUMFX: op = OP_ADD; src = R_FX; break;
    case S_SUMFY: op = OP_ADD; src = R_FY; break;
    case S_SUMFZ: op = OP_ADD; src = R_FZ; break;
    case S_MAXFZ: op = OP_MAX; src = -2; break;
    case S_SUMT: op = OP_ADD; src = R_TQ; break;
    case S_MAXT: op = OP_MAX; src = R_TQ; break;
    case S_SUMX: op = OP_ADD; src = R_X; break;
    case S_SUMY: op = OP_ADD; src = R_Y; break;
    case S_MINX: op = OP_MIN; src = R_X; break;
    case S_MAXX: op = OP_MAX; src = R_X; break;
    case S_MINY: op = OP_MIN; src = R_Y; break;
    case S_MAXY: op = OP_MAX; src = R_Y; break;
    case S_SUMZC: op = OP_ADD; src = R_ZC; cond = 1; break;
    case S_WZC: op = OP_ADD; src = R_FXY; cond = 1; break;
    case S_SUMW: op = OP_ADD; src = R_W; cond = 2; break;
    case S_MAXW: op = OP_MAX; src = -3; cond = 2; break;
    case S_ENGT: op = OP_ADD; src = R_ENG; break;
    default: break;
  }
  if (cond == 2 && !wallOn) op = OP_NONE;
  // the record rows of this lane's operand and condition
  const int srow = src >= 0 ? src : src == -1 ? R_ONE : src == -2 ? R_NFZ : R_AW;
  const int crow = cond == 1 ? R_ZCOK : R_ONE;
  rec[R_ONE][lane] = 1.0;
  double stv = op != OP_NONE ? E.st[lane] : 0.0;
  int64_t hist_i = E.hist_i, hist_count = E.hist_count, ts_i = E.ts_i, ts_count = E.ts_count;
  int64_t ws_i = E.ws_i, ws_count = E.ws_count, wsp_i = E.wsp_i, wsp_count = E.wsp_count;
  int64_t nev = 0;
  const int64_t ev_cap = E.ev_cap;
  // steps per chunk: at most a tooth period (no delay-line cell twice), 64 and GCAP entries
  int chMax = spt < 64 ? spt : 64;
  if (NS > 0 && GCAP / NS < chMax) chMax = GCAP / NS;
  // this lane's first visit: its tooth and its slice (or end ring), the same every step
  const int iL = per > 0 ? lane / per : 0, rL = per > 0 ? lane - iL * per : 0;
  // the sum lanes' and the tooth lanes' rows of the visit terms
  const double* trow = lane < T_SUMS ? term[lane] : term[T_VOL];
  wave_sync();

  for (int64_t s0 = 0; s0 < nSteps; s0 += chMax) {
    const int ch = nSteps - s0 < chMax ? (int)(nSteps - s0) : chMax;
    const int prc = pr;
    // ---- the chunk's edge angles and delay-line cells
    for (int e = lane; e < ch * NS; e += 64) {
      const int sc = e / NS, slot = e - sc * NS;
      const int i = slot / per, r = slot - i * per;
      int prs = prc + sc;
      if (prs >= spr) prs -= spr;
      const double theta = prs * dphi;
      double phi = r < nz ? mcm::fmod_2pi((theta + i * pitch) - L.sLag[r]) : mcm::fmod_2pi(theta + i * pitch);
      if (phi < 0) phi += TWO_PI;
      double sn, cs;
      mcm::sincos_js(phi, sn, cs);
      L.gphi[e] = phi; L.gsn[e] = sn; L.gcs[e] = cs;
      if (r < nz) {
        int pi = prs + i * spt;
        if (pi >= spr) pi -= spr;
        L.win[e] = surf[(size_t)pi * MSL + r];
      }
    }
    wave_sync();
    for (int sc = 0; sc < ch; sc++) {
      const double rx = x - w * wnx, ry = y - w * wny;
      const double xm = rx * 1e3, ym = ry * 1e3;
      const double rvx = vxs - wv * wnx, rvy = vys - wv * wny;
      // ---- visits (src/cutting.js:172-237)
      for (int slot = lane; slot < NS; slot += 64) {
        const int e = sc * NS + slot;
        const int i = slot == lane ? iL : slot / per, r = slot == lane ? rL : slot - (slot / per) * per;
        double a[T_COUNT];
#pragma unroll
        for (int q = 0; q < T_COUNT; q++) a[q] = 0.0;
        double ft = 0, fr = 0;
        int eng = 0;
        const double phi = L.gphi[e], sn = L.gsn[e], cs = L.gcs[e];
        if (r < nz) {
          const int k = r;
          const double nxg = sn * ux + cs * vx, nyg = sn * uy + cs * vy;
          const double rn = xm * nxg + ym * nyg + L.dR[i];
          const double prev = (double)L.win[e];
          const double feedH = fz * sn * L.sKap[k];
          const int bin = (int)(phi * binScale);
          const double hb = bin < MC_BINS ? (double)L.Hb[bin] : __builtin_nan("");
          if (hb <= L.sZ[k] || sn <= 0) {
            L.win[e] = (float)rn;
          } else {
            double h = feedH + rn - prev;
            if (h <= 0) {
              L.win[e] = (float)(prev - feedH);
            } else {
              L.win[e] = (float)rn;
              if (h > hMax) h = hMax;
              const double xs = H_REF / (h > H_MIN ? h : H_MIN);
              const double se = !mcOn ? 1.0 : cellPow ? mcm::pow_cell(xs, mc, my1, myr) : pow_js_outlined(xs, mc);
              const double bk = L.sB[k], sr = L.sR[k], sz = L.sZ[k];
              const double Ft = Ktc * se * h * bk + L.cTe[k];
              double Fr = Krc * se * h * bk + L.cRe[k];
              const double Fa = Kac * se * h * bk + L.cAe[k];
              if (pdOn) Fr += L.cPd[k] * (rvx * nxg + rvy * nyg);
              const double txg = cs * ux - sn * vx, tyg = cs * uy - sn * vy;
              a[T_FX] = -Ft * txg - Fr * nxg;
              a[T_FY] = -Ft * tyg - Fr * nyg;
              a[T_FZ] = -Fa;
              a[T_TQ] = Ft * sr;
              a[T_HB] = h * bk;
              a[T_B] = bk;
              a[T_RB] = sr * h * bk;
              a[T_VOL] = h * bk * sr * dphi;
              a[T_PHI] = phi;
              a[T_Z] = sz;
              ft = Ft; fr = Fr;
              eng = 1;
            }
          }
        } else {
          const int j = r - nz;
          const int rb = (int)(phi / TWO_PI * EBINS);
          const double depth = rb < EBINS ? (double)L.Eb[rb * ERINGS + j] : __builtin_nan("");
          if (!(depth <= 2e-4)) {
            const double txg = cs * ux - sn * vx, tyg = cs * uy - sn * vy;
            const double c = L.ring[ERINGS + j], ds = L.ring[3 * ERINGS + j], rr = L.ring[j], rz = L.ring[2 * ERINGS + j];
            const double h = (depth < fza ? depth : fza) * c;
            const double xs = H_REF / (h > H_MIN ? h : H_MIN);
            const double se = !mcOn ? 1.0 : cellPow ? mcm::pow_cell(xs, mc, my1, myr) : pow_js_outlined(xs, mc);
            const double Ft = Ktc * se * h * ds + L.cTe[MS + j];
            const double Fn = Krc * se * h * ds + L.cRe[MS + j];
            a[T_FX] = -Ft * txg;
            a[T_FY] = -Ft * tyg;
            a[T_FZ] = -(Fn * c);
            a[T_TQ] = Ft * rr;
            a[T_HB] = h * ds;
            a[T_B] = ds;
            a[T_RB] = rr * h * ds;
            a[T_VOL] = h * ds * rr * dphi;
            a[T_PHI] = phi;
            a[T_Z] = tipZ + rz;
            ft = Ft;
            eng = 1;
          }
        }
#pragma unroll
        for (int q = 0; q < T_COUNT; q++) term[q][slot] = a[q];
        teng[slot] = eng;
        L.dFt[e] = ft; L.dFr[e] = fr;
      }
      wave_sync();
      // ---- serial sums in the reference's visit order (sum lanes), and a tooth lane's chip
      // volume (src/cutting.js:213) and last engaged visit over its own visits, eight terms
      // loaded at a time (the sums start at +0 or at a non-negative carried sum, so they are
      // never -0 and the +0 padding past the last visit leaves them unchanged)
      double acc = lane < T_HB ? 0.0 : lane < T_SUMS ? chain : isTooth ? tvol : 0.0;
      int last = -1;
      const int qa = isTooth ? tooth * per : 0, qb = isTooth ? qa + per : NS;
      for (int q0 = qa; q0 < qb; q0 += 8) {
        double v[8];
        int ev[8];
#pragma unroll
        for (int u = 0; u < 8; u++) { v[u] = q0 + u < qb ? trow[q0 + u] : 0.0; ev[u] = q0 + u < qb ? teng[q0 + u] : 0; }
#pragma unroll
        for (int u = 0; u < 8; u++) { acc += v[u]; last = ev[u] ? q0 + u : last; }
      }
      if (lane < T_HB) res[lane] = acc;
      else if (lane < T_SUMS) chain = acc;
      bool push = false;
      double pvol = 0, pphi = 0, pz = 0;
      bool engagedNow = false;
      if (isTooth) {
        tvol = acc;
        engagedNow = last >= 0;
        if (engagedNow) {
          tphi = term[T_PHI][last];
          tz = term[T_Z][last];
        } else if (tEng) {
          if (tvol > 0) { push = true; pvol = tvol; pphi = tphi; pz = tz; }
          tvol = 0;
        }
        tEng = engagedNow ? 1 : 0;
      }
      // chip events in tooth order (src/cutting.js:238-241)
      const unsigned long long pmask = __ballot(push);
      const unsigned long long emask = __ballot(engagedNow);
      if (push) {
        const int64_t at = nev + __popcll(pmask & ((1ull << lane) - 1));
        if (at < ev_cap) {
          double* q = E.ev + 3 * at;
          q[0] = pvol; q[1] = pphi; q[2] = pz;
        }
      }
      nev += __popcll(pmask);
      wave_sync();
      const double Fx = res[T_FX], Fy = res[T_FY];
      // ---- modal integration, semi-implicit Euler (src/cutting.js:246-255): each mode on its
      // lane, the displacements and velocities summed in mode order
      if (xmode || ymode) {
        mv += ((lane < MC_MAX_MODES ? Fx : Fy) - mcd * mv - mk * mq) / mmass * dt;
        mq += mv * dt;
      }
      {
        // every mode's displacement and velocity read from its lane, then summed from +0 in mode
        // order (+0 for the modes past nmx / nmy: x starts at +0 and a sum of +0 with anything
        // is never -0, so the padding leaves it unchanged)
        double q[2 * MC_MAX_MODES], v[2 * MC_MAX_MODES];
#pragma unroll
        for (int j = 0; j < 2 * MC_MAX_MODES; j++) { q[j] = read_lane(mq, j); v[j] = read_lane(mv, j); }
        x = 0; y = 0; vxs = 0; vys = 0;
#pragma unroll
        for (int j = 0; j < MC_MAX_MODES; j++) {
          x += j < nmx ? q[j] : 0.0; vxs += j < nmx ? v[j] : 0.0;
          y += j < nmy ? q[MC_MAX_MODES + j] : 0.0; vys += j < nmy ? v[MC_MAX_MODES + j] : 0.0;
        }
      }
      // ---- wall mode (src/cutting.js:256-264)
      if (wallOn) {
        const double Fw = -(Fx * wnx + Fy * wny);
        wv += (Fw - wc * wv - wk * w) / wm * dt;
        w += wv * dt;
        if (w > 5e-4) { w = 5e-4; if (wv > 0) wv = 0; }
        else if (w < -2e-4) { w = -2e-4; if (wv < 0) wv = 0; }
      }
      // ---- the step's record
      if (lane == 0) {
        rec[R_FX][sc] = Fx; rec[R_FY][sc] = Fy; rec[R_FZ][sc] = res[T_FZ]; rec[R_TQ][sc] = res[T_TQ];
        rec[R_X][sc] = x; rec[R_Y][sc] = y; rec[R_W][sc] = w; rec[R_ENG][sc] = (double)__popcll(emask);
        rec[R_NFZ][sc] = -res[T_FZ]; rec[R_AW][sc] = __builtin_fabs(w);
      }
    }
    wave_sync();
    // ---- after the chunk: the force magnitudes of every visit, |(Ft, Fr)| on a slice and Ft
    // on an end ring, with their moments about the tip (src/cutting.js:207-208,230), four
    // visits per lane at a time
    for (int e0 = lane; e0 < ch * NS; e0 += 4 * 64) {
      double fs[4], fm[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + u * 64;
        const int ee = e < ch * NS ? e : e0;
        const int slot = ee % NS, r = slot % per;
        const double a = L.dFt[ee], b = L.dFr[ee];
        fs[u] = r < nz ? hypot_finite(a, b) : a;
        fm[u] = r < nz ? L.sZ[r] - tipZ : L.ring[2 * ERINGS + r - nz];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + u * 64;
        if (e < ch * NS) { L.dFt[e] = fs[u]; L.dFr[e] = fs[u] * fm[u]; }
      }
    }
    wave_sync();
    // lane sc: the step's Fsum and Fr2z summed in visit order from +0, |(Fx, Fy)|, and the
    // moment centre term of the statistics; and whether the step takes a sample of the rings
    // or of the history
    {
      const bool isStep = lane < ch;
      const int sl = isStep ? lane : 0;
      double fsum = 0, fr2z = 0;
      for (int q0 = 0; q0 < NS; q0 += 8) {
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          a[u] = q0 + u < NS ? L.dFt[sl * NS + q0 + u] : 0.0;
          b[u] = q0 + u < NS ? L.dFr[sl * NS + q0 + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) { fsum += a[u]; fr2z += b[u]; }
      }
      if (isStep) {
        const double Fxy = hypot_finite(rec[R_FX][sl], rec[R_FY][sl]);
        rec[R_FXY][sl] = Fxy;
        rec[R_ZCOK][sl] = fsum > 0 ? 1.0 : 0.0;
        rec[R_ZC][sl] = fsum > 0 ? fr2z / fsum * Fxy : 0.0;
      }
    }
    wave_sync();
    // the statistics (src/cutting.js:279-287), lane f over the chunk's steps in order, eight
    // records loaded at a time
    // (selects only: the lanes' operations differ, and a branch per operation serialised them)
    for (int s8 = 0; s8 < ch; s8 += 8) {
      double o[8], fxy[8], c[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int sc = s8 + u < ch ? s8 + u : s8;
        fxy[u] = rec[R_FXY][sc];
        o[u] = rec[srow][sc];
        c[u] = rec[crow][sc];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const bool in = s8 + u < ch;
        runMaxF = in ? jmax_sel(runMaxF, fxy[u]) : runMaxF;
        const double nv = op == OP_ADD ? stv + o[u] : op == OP_MAX ? jmax_sel(stv, o[u]) : jmin_sel(stv, o[u]);
        stv = (in && op != OP_NONE && c[u] != 0) ? nv : stv;
      }
    }
    // the rings and the history: step sc's pr, its phase in the tooth period and its count;
    // the samples are numbered in step order with ballots and written by their steps' lanes
    {
      const bool isStep = lane < ch;
      const int sl = isStep ? lane : 0;
      int prs = prc + sl + 1;
      if (prs >= spr) prs -= spr;
      const int phs = prs % spt, tqs = prs / spt;
      const bool tsS = isStep && phs == 0, wpS = isStep && phs == piStep;
      const bool hiS = isStep && sl + 1 >= dleft && (sl + 1 - dleft) % decim == 0;
      const unsigned long long mTs = __ballot(tsS), mWp = __ballot(wpS), mHi = __ballot(hiS);
      const unsigned long long below = (1ull << lane) - 1;
      const double xs = rec[R_X][sl], ys = rec[R_Y][sl], ws = rec[R_W][sl];
      if (tsS) {
        // the tooth now at phi = 0 generates the +v wall (src/cutting.js:265-272)
        const double d0 = L.dR[(N - tqs) % N] * 1e-3;
        const int64_t a = (ts_i + __popcll(mTs & below)) % RN, b = (ws_i + __popcll(mTs & below)) % RN;
        E.rings[a] = (float)(xs - ws * wnx);
        E.rings[RN + a] = (float)(ys - ws * wny);
        E.rings[2 * RN + b] = (float)(xs - ws * wnx + d0 * vx);
        E.rings[3 * RN + b] = (float)(ys - ws * wny + d0 * vy);
      }
      if (wpS) {
        const long long qn = (long long)mcm::round_js((JS_PI - prs * dphi) / pitch);
        const int ip = (int)(((qn % N) + N) % N);
        const double dp = L.dR[ip] * 1e-3;
        const int64_t a = (wsp_i + __popcll(mWp & below)) % RN;
        E.rings[4 * RN + a] = (float)(xs - ws * wnx - dp * vx);
        E.rings[5 * RN + a] = (float)(ys - ws * wny - dp * vy);
      }
      if (hiS) {
        const int64_t a = (hist_i + __popcll(mHi & below)) % HN;
        E.hist[a] = (float)rec[R_FX][sl];
        E.hist[HN + a] = (float)rec[R_FY][sl];
        E.hist[2 * HN + a] = (float)rec[R_FZ][sl];
        E.hist[3 * HN + a] = (float)((xs - ws * wnx) * 1e6);
        E.hist[4 * HN + a] = (float)((ys - ws * wny) * 1e6);
      }
      const int nts = __popcll(mTs), nwp = __popcll(mWp), nhi = __popcll(mHi);
      ts_i = (ts_i + nts) % RN; ts_count += nts;
      ws_i = (ws_i + nts) % RN; ws_count += nts;
      wsp_i = (wsp_i + nwp) % RN; wsp_count += nwp;
      hist_i = (hist_i + nhi) % HN; hist_count += nhi;
      dleft = nhi ? decim - (ch - (dleft + (nhi - 1) * decim)) : dleft - ch;
      stepCount += ch;
      pr = prc + ch;
      if (pr >= spr) pr -= spr;
      ph = pr % spt; tq = pr / spt;
    }
    // the time and the mean wall deflection, summed step by step
    for (int s8 = 0; s8 < ch; s8 += 8) {
      double wl[8];
#pragma unroll
      for (int u = 0; u < 8; u++) wl[u] = rec[R_W][s8 + u < ch ? s8 + u : s8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const bool in = s8 + u < ch;
        t = in ? t + dt : t;
        wSum = in && wallOn ? wSum + wl[u] : wSum;
      }
    }
    // ---- the chunk's delay-line cells back to the engine's record
    for (int e = lane; e < ch * NS; e += 64) {
      const int sc = e / NS, slot = e - sc * NS;
      const int i = slot / per, r = slot - i * per;
      if (r < nz) {
        int prs = prc + sc;
        if (prs >= spr) prs -= spr;
        int pi = prs + i * spt;
        if (pi >= spr) pi -= spr;
        surf[(size_t)pi * MSL + r] = L.win[e];
      }
    }
    wave_sync();
  }
  // ---- write the state back
  if (lane == T_HB) E.st[S_SUMHB] = chain;
  if (lane == T_B) E.st[S_SUMB] = chain;
  if (lane == T_RB) E.st[S_SUMRB] = chain;
  if (op != OP_NONE) E.st[lane] = stv;
  if (isTooth) {
    E.toothVol[tooth] = tvol; E.toothPhi[tooth] = tphi; E.toothZ[tooth] = tz; E.toothEngaged[tooth] = tEng;
  }
  if (xmode) { E.qx[mj] = mq; E.vxm[mj] = mv; }
  if (ymode) { E.qy[mj] = mq; E.vym[mj] = mv; }
  if (lane == 0) {
    E.x = E.x0 = x; E.y = E.y0 = y; E.vx = E.vx0 = vxs; E.vy = E.vy0 = vys;
    E.t = t; E.pr = pr; E.w = w; E.wv = wv;
    E.runMeanW = nSteps ? wSum / (double)nSteps : 0.0;
    E.stepCount = stepCount;
    E.hist_i = hist_i; E.hist_count = hist_count; E.ts_i = ts_i; E.ts_count = ts_count;
    E.ws_i = ws_i; E.ws_count = ws_count; E.wsp_i = wsp_i; E.wsp_count = wsp_count;
    E.nev = nev;
    E.runMaxF = runMaxF;
  }
  wave_sync();
}

// setWork(null), src/cutting.js:89
__host__ __device__ __forceinline__ void eng_set_work_null(EngDev& h) {
  h.runMeanW = 0; h.wk = 0; h.wm = 1; h.wc = 0; h.w = 0; h.wv = 0; h.wnx = 1; h.wny = 0;
}

// setWork(wall), src/cutting.js:90-95: the wall state carries over while the normal stays
// within about 37 degrees of the previous one
__host__ __device__ __forceinline__ void eng_set_work(EngDev& h, double k, double m, double zeta, double nx, double ny) {
  if (truthy(h.wk) && nx * h.wnx + ny * h.wny < 0.8) { h.w = 0; h.wv = 0; }
  h.wk = k * 1e3; h.wm = m; h.wc = 2 * zeta * __builtin_sqrt(h.wk * m);
  h.wnx = nx; h.wny = ny;
}

// the simulator's tool shapes as the engine and the stamps see them: flat, ball, or bull-nose
// with corner radius rc and its tip table (lut of lutn heights every MC_LUT_DR)
struct ToolShape {
  int ball;           // the engine's ball rule and the stamp's sphere
  double Rt;          // tool radius
  double rc;          // bull-nose corner radius (> 0: the profile rAt below)
  const float* lut;   // tip table or null
  int lutn;
};

// an imported cutting-edge envelope: n >= 2 radii dz apart from the tip (null r: none)
struct EdgeProfile {
  const double* r;
  int n;
  double dz;
};

// sampleEdge (src/assembly.js:31-35): the envelope's radius at height z, linear between
// the samples, the end samples held beyond the table
__host__ __device__ __forceinline__ double edge_rat(const EdgeProfile& e, double z) {
  const double f = z / e.dz;
  const double i = jmin(e.n - 2, jmax(0, __builtin_floor(f)));
  const double u = jmin(1, jmax(0, f - i));
  const int k = (int)i;
  return e.r[k] * (1 - u) + e.r[k + 1] * u;
}

// cutterRadiusAt (src/assembly.js:23-30) of a bull-nose tool
__host__ __device__ __forceinline__ double bull_rat(double R, double rc, double z) {
  if (rc > 0 && z < rc) {
    const double q = rc - z;
    return R - rc + __builtin_sqrt(jmax(0, rc * rc - q * q));
  }
  return R;
}

// setEngagement(u, fz, tipZ, ap), src/cutting.js:98-124: the flat / ball slices, the caller's
// profile radii sR / sK (tool.rAt), or an imported edge (edge) or a bull-nose profile (rc > 0)
// evaluated here
__host__ __device__ __forceinline__ int eng_set_engagement(EngDev& h, double ux, double uy, double fz, double tipZ,
                                                          double ap, const double* sR, const double* sK, bool ball,
                                                          double rc = 0, const EdgeProfile* edge = nullptr) {
  h.ux = ux; h.uy = uy; h.ex = -uy; h.ey = ux;
  h.fz = fz; h.tipZ = tipZ;
  const int nz = ap <= 0 ? 0 : (int)jmin(h.maxSlices, __builtin_ceil(ap / h.dz));
  h.nz = nz;
  h.db = h.dz;
  for (int k = 0; k < nz; k++) {
    h.sB[k] = jmin(h.dz, ap - k * h.dz);
    const double zr = k * h.dz + 0.5 * h.sB[k];
    h.sZ[k] = tipZ + zr;
    double r = h.R, kap = 1;
    if (sR) { r = sR[k]; kap = sK[k]; }
    else if (edge && edge->r) {
      r = jmax(1e-3, edge_rat(*edge, zr));
      const double dr = (edge_rat(*edge, zr + 0.05) - edge_rat(*edge, jmax(0, zr - 0.05))) / (zr < 0.05 ? 0.05 + zr : 0.1);
      kap = 1 / __builtin_sqrt(1 + dr * dr);
    } else if (rc > 0) {
      r = jmax(1e-3, bull_rat(h.R, rc, zr));
      const double dr = (bull_rat(h.R, rc, zr + 0.05) - bull_rat(h.R, rc, jmax(0, zr - 0.05))) / (zr < 0.05 ? 0.05 + zr : 0.1);
      kap = 1 / __builtin_sqrt(1 + dr * dr);
    } else if (ball && zr < h.R) {
      const double q = h.R - zr;
      r = __builtin_sqrt(jmax(1e-6, h.R * h.R - q * q));
      kap = r / h.R;
    }
    h.sR[k] = r; h.sKap[k] = kap;
    h.sLag[k] = zr * h.tanB / h.R;
  }
  return nz;
}

// the end-face shape zAt of the simulator's tools (src/sim.js:216-218): the tip table when
// there is one, the sphere for a ball, 0 for a flat end
__host__ __device__ __forceinline__ double eng_zat(const ToolShape& t, double r) {
  if (t.lut) {
    const double i = mcm::round_js(r / MC_LUT_DR);
    return (double)t.lut[(int)jmin(t.lutn - 1, i)];
  }
  return t.ball ? t.Rt - __builtin_sqrt(jmax(0, t.Rt * t.Rt - r * r)) : 0.0;
}

// setEndEngagement(fza, rOut, zAt), src/cutting.js:130-143: rings from the caller (R, cos,
// z, ds) or from zAt = eng_zat(shape, .)
__host__ __device__ __forceinline__ void eng_set_end_engagement(EngDev& h, double fza, double rOut,
                                                               const double* rings, const ToolShape& t) {
  h.fza = fza > 0 ? fza : 0;
  h.endR = rOut;
  const double dr = rOut / ERINGS;
  for (int j = 0; j < ERINGS; j++) {
    if (rings) {
      h.ringR[j] = rings[j]; h.ringCos[j] = rings[ERINGS + j];
      h.ringZ[j] = rings[2 * ERINGS + j]; h.ringDs[j] = rings[3 * ERINGS + j];
    } else {
      const double r = (j + 0.5) * dr;
      const double slope = (eng_zat(t, jmin(rOut, r + 0.25 * dr)) - eng_zat(t, jmax(0, r - 0.25 * dr))) / (0.5 * dr);
      const double c = jmax(0.22, 1 / __builtin_sqrt(1 + slope * slope));
      h.ringR[j] = r; h.ringCos[j] = c; h.ringDs[j] = dr / c;
      h.ringZ[j] = eng_zat(t, r);
    }
  }
}

// the constants run() derives from the material, the edge coefficients and the
// engagement (src/cutting.js:146-165)
__host__ __device__ __forceinline__ void eng_run_constants(EngDev& h, const mc_material& m, bool nonlinear, double Kte,
                                                          double Kre) {
  h.Ktc = m.Ktc; h.Krc = m.Krc; h.Kac = m.Kac; h.Kae = m.Kae;
  h.Kte = Kte; h.Kre = Kre;
  h.mc = nonlinear ? m.mc : 0.0;
  h.pdK = nonlinear ? m.Ci * h.db * 1e-3 / (h.omega * 1e-3) : 0.0;
  const double TWO_PI = 2 * JS_PI;
  h.binScale = MC_BINS / TWO_PI;
  h.pitch = TWO_PI / h.N;
  h.dphi = h.omega * h.dt;
  h.hMax = jmax(0.1, 3 * h.fz);
}

}  // namespace mce
