// engine.hip — a device-resident CuttingEngine (src/cutting.js:21-326) for gfx950.
//
// The engine object a simulator keeps across its sub-steps lives in device memory: modal
// state, delay line (surf), wall mode, per-tooth chip volumes, statistics, the decimated
// history and the tooth / wall sample rings.  A run of n steps is one workgroup of one
// wavefront per engine; mc_engine_run_batch puts many engines in one launch.
//
// Per step, the lanes split the (tooth, slice) side-edge visits and the (tooth, ring)
// end-edge visits; each lane evaluates its visits' engagement, chip thickness and forces,
// writes the delay line and leaves every quantity the reference accumulates in LDS, at the
// visit's position in the reference's loop order (tooth by tooth, its slices, then its end
// rings).  Nine lanes then each sum one of the accumulations (Fx, Fy, Fz, torque, force
// sum, its moment, and the three engagement statistics) serially in that order, and one
// lane per tooth sums the tooth's chip volume and decides its chip event, so every FP64
// sum is the reference's, bit for bit.  The modal and wall updates run redundantly on all
// lanes; lane 0 keeps the step statistics and writes the sample rings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_host.h"
#include "mc_math.h"

namespace {

constexpr int EBINS = MC_END_BINS, ERINGS = MC_END_RINGS;
constexpr int HN = MC_ENGINE_HIST, RN = MC_ENGINE_RING;
constexpr int MS = MC_MAX_SLICES;
constexpr int SLOT_CAP = MC_MAX_FLUTES * (MS + ERINGS);
constexpr double JS_PI = 3.141592653589793;
constexpr double H_REF = 0.05, H_MIN = 0.002;

enum { S_N, S_SUMF, S_MAXF, S_SUMFX, S_SUMFY, S_SUMFZ, S_MAXFZ, S_SUMT, S_MAXT, S_SUMX, S_SUMY, S_MINX,
       S_MAXX, S_MINY, S_MAXY, S_SUMZC, S_WZC, S_SUMW, S_MAXW, S_SUMHB, S_SUMB, S_SUMRB, S_ENGT, S_COUNT };

// accumulations kept per visit (the order of the chain lanes)
enum { T_FX, T_FY, T_FZ, T_TQ, T_FS, T_FR2Z, T_HB, T_B, T_RB, T_VOL, T_PHI, T_Z, T_COUNT };

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
  float* surf;   // [spr][maxSlices]
  float* hist;   // [5][HN]: Fx Fy Fz x y
  float* rings;  // [6][RN]: ts x, ts y, ws x, ws y, wsp x, wsp y
  double* ev;    // [ev_cap][3]: vol, phi, z
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
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

__global__ __launch_bounds__(64) void engine_run(EngDev* const* engines) {
  EngDev& E = *engines[blockIdx.x];
  __shared__ double term[T_COUNT][SLOT_CAP];
  __shared__ int teng[SLOT_CAP];
  __shared__ double res[8];
  const int lane = threadIdx.x;
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
  const int64_t nSteps = E.nsteps;
  float* const surf = E.surf;

  // modal state, redundantly on every lane; a state set from outside enters mode 0
  double qx[MC_MAX_MODES], vxm[MC_MAX_MODES], qy[MC_MAX_MODES], vym[MC_MAX_MODES];
#pragma unroll
  for (int j = 0; j < MC_MAX_MODES; j++) { qx[j] = E.qx[j]; vxm[j] = E.vxm[j]; qy[j] = E.qy[j]; vym[j] = E.vym[j]; }
  qx[0] += E.x - E.x0; qy[0] += E.y - E.y0;
  vxm[0] += E.vx - E.vx0; vym[0] += E.vy - E.vy0;
  double x = E.x, y = E.y, vxs = E.vx, vys = E.vy, t = E.t, w = E.w, wv = E.wv, wSum = 0;
  int pr = E.pr;
  int64_t stepCount = E.stepCount;
  // chain lanes 6..8 carry the engagement statistics across steps; tooth lanes their tooth
  const int tooth = lane - 16;
  const bool isTooth = tooth >= 0 && tooth < N;
  double chain = lane == 6 ? E.st[S_SUMHB] : lane == 7 ? E.st[S_SUMB] : lane == 8 ? E.st[S_SUMRB] : 0.0;
  double tvol = isTooth ? E.toothVol[tooth] : 0.0, tphi = isTooth ? E.toothPhi[tooth] : 0.0,
         tz = isTooth ? E.toothZ[tooth] : 0.0;
  int tEng = isTooth ? E.toothEngaged[tooth] : 0;
  // lane 0: the step statistics, rings and events
  double st[S_COUNT];
#pragma unroll
  for (int k = 0; k < S_COUNT; k++) st[k] = E.st[k];
  int64_t hist_i = E.hist_i, hist_count = E.hist_count, ts_i = E.ts_i, ts_count = E.ts_count;
  int64_t ws_i = E.ws_i, ws_count = E.ws_count, wsp_i = E.wsp_i, wsp_count = E.wsp_count;
  int64_t nev = 0;
  const int64_t ev_cap = E.ev_cap;

  for (int64_t s = 0; s < nSteps; s++) {
    const double theta = pr * dphi;
    const double rx = x - w * wnx, ry = y - w * wny;
    const double xm = rx * 1e3, ym = ry * 1e3;
    const double rvx = vxs - wv * wnx, rvy = vys - wv * wny;
    // ---- visits (src/cutting.js:172-237)
    for (int slot = lane; slot < NS; slot += 64) {
      const int i = slot / per, r = slot - i * per;
      double a[T_COUNT];
#pragma unroll
      for (int q = 0; q < T_COUNT; q++) a[q] = 0.0;
      int eng = 0;
      if (r < nz) {
        const int k = r;
        double phi = mcm::fmod_2pi((theta + i * pitch) - E.sLag[k]);
        if (phi < 0) phi += TWO_PI;
        double sn, cs;
        mcm::sincos_js(phi, sn, cs);
        const double nxg = sn * ux + cs * vx, nyg = sn * uy + cs * vy;
        const double rn = xm * nxg + ym * nyg + E.dR[i];
        int pi = pr + i * spt;
        if (pi >= spr) pi -= spr;
        float* const cell = surf + (size_t)pi * MSL + k;
        const double prev = (double)*cell;
        const double feedH = fz * sn * E.sKap[k];
        const int bin = (int)(phi * binScale);
        const double hb = bin < MC_BINS ? (double)E.Hb[bin] : __builtin_nan("");
        if (hb <= E.sZ[k] || sn <= 0) {
          *cell = (float)rn;
        } else {
          double h = feedH + rn - prev;
          if (h <= 0) {
            *cell = (float)(prev - feedH);
          } else {
            *cell = (float)rn;
            if (h > hMax) h = hMax;
            const double se = mcOn ? mcm::pow_js(H_REF / (h > H_MIN ? h : H_MIN), mc) : 1.0;
            const double bk = E.sB[k], sr = E.sR[k], sz = E.sZ[k];
            const double Ft = Ktc * se * h * bk + Kte * bk;
            double Fr = Krc * se * h * bk + Kre * bk;
            const double Fa = Kac * se * h * bk + Kae * bk;
            if (pdOn) Fr += pdK * (bk / db) / sr * (rvx * nxg + rvy * nyg);
            const double txg = cs * ux - sn * vx, tyg = cs * uy - sn * vy;
            a[T_FX] = -Ft * txg - Fr * nxg;
            a[T_FY] = -Ft * tyg - Fr * nyg;
            a[T_FZ] = -Fa;
            a[T_TQ] = Ft * sr;
            const double fr = mcm::hypot_js(Ft, Fr);
            a[T_FS] = fr;
            a[T_FR2Z] = fr * (sz - tipZ);
            a[T_HB] = h * bk;
            a[T_B] = bk;
            a[T_RB] = sr * h * bk;
            a[T_VOL] = h * bk * sr * dphi;
            a[T_PHI] = phi;
            a[T_Z] = sz;
            eng = 1;
          }
        }
      } else {
        const int j = r - nz;
        double phi = mcm::fmod_2pi(theta + i * pitch);
        if (phi < 0) phi += TWO_PI;
        const int rb = (int)(phi / TWO_PI * EBINS);
        const double depth = rb < EBINS ? (double)E.Eb[rb * ERINGS + j] : __builtin_nan("");
        if (!(depth <= 2e-4)) {
          double sn, cs;
          mcm::sincos_js(phi, sn, cs);
          const double txg = cs * ux - sn * vx, tyg = cs * uy - sn * vy;
          const double c = E.ringCos[j], ds = E.ringDs[j], rr = E.ringR[j], rz = E.ringZ[j];
          const double h = (depth < fza ? depth : fza) * c;
          const double se = mcOn ? mcm::pow_js(H_REF / (h > H_MIN ? h : H_MIN), mc) : 1.0;
          const double Ft = Ktc * se * h * ds + Kte * ds;
          const double Fn = Krc * se * h * ds + Kre * ds;
          a[T_FX] = -Ft * txg;
          a[T_FY] = -Ft * tyg;
          a[T_FZ] = -(Fn * c);
          a[T_TQ] = Ft * rr;
          a[T_FS] = Ft;
          a[T_FR2Z] = Ft * rz;
          a[T_HB] = h * ds;
          a[T_B] = ds;
          a[T_RB] = rr * h * ds;
          a[T_VOL] = h * ds * rr * dphi;
          a[T_PHI] = phi;
          a[T_Z] = tipZ + rz;
          eng = 1;
        }
      }
#pragma unroll
      for (int q = 0; q < T_COUNT; q++) term[q][slot] = a[q];
      teng[slot] = eng;
    }
    wave_sync();
    // ---- serial sums in the reference's visit order
    if (lane < 9) {
      double acc = lane < 6 ? 0.0 : chain;
      for (int q = 0; q < NS; q++) acc += term[lane][q];
      if (lane < 6) res[lane] = acc; else chain = acc;
    }
    bool push = false;
    double pvol = 0, pphi = 0, pz = 0;
    bool engagedNow = false;
    if (isTooth) {
      int last = -1;
      for (int r = 0; r < per; r++) {
        const int q = tooth * per + r;
        tvol += term[T_VOL][q];
        if (teng[q]) last = q;
      }
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
    const double Fx = res[T_FX], Fy = res[T_FY], Fz = res[T_FZ], Tq = res[T_TQ], Fsum = res[T_FS],
                 Fr2z = res[T_FR2Z];
    // ---- modal integration, semi-implicit Euler (src/cutting.js:246-255)
    x = 0; y = 0; vxs = 0; vys = 0;
#pragma unroll
    for (int j = 0; j < MC_MAX_MODES; j++) {
      if (j < nmx) {
        vxm[j] += (Fx - E.cx[j] * vxm[j] - E.kx[j] * qx[j]) / E.mx[j] * dt;
        qx[j] += vxm[j] * dt;
        x += qx[j];
        vxs += vxm[j];
      }
    }
#pragma unroll
    for (int j = 0; j < MC_MAX_MODES; j++) {
      if (j < nmy) {
        vym[j] += (Fy - E.cy[j] * vym[j] - E.ky[j] * qy[j]) / E.my[j] * dt;
        qy[j] += vym[j] * dt;
        y += qy[j];
        vys += vym[j];
      }
    }
    // ---- wall mode (src/cutting.js:256-264)
    if (wallOn) {
      const double Fw = -(Fx * wnx + Fy * wny);
      wv += (Fw - wc * wv - wk * w) / wm * dt;
      w += wv * dt;
      if (w > 5e-4) { w = 5e-4; if (wv > 0) wv = 0; }
      else if (w < -2e-4) { w = -2e-4; if (wv < 0) wv = 0; }
      st[S_SUMW] += w;
      st[S_MAXW] = jmax(st[S_MAXW], __builtin_fabs(w));
      wSum += w;
    }
    if (++pr == spr) pr = 0;
    const int ph = pr % spt;
    if (ph == 0) {
      // the tooth now at phi = 0 generates the +v wall (src/cutting.js:265-272)
      const double d0 = E.dR[(N - (pr / spt) % N) % N] * 1e-3;
      if (lane == 0) {
        E.rings[ts_i] = (float)(x - w * wnx);
        E.rings[RN + ts_i] = (float)(y - w * wny);
        E.rings[2 * RN + ws_i] = (float)(x - w * wnx + d0 * vx);
        E.rings[3 * RN + ws_i] = (float)(y - w * wny + d0 * vy);
      }
      ts_i = (ts_i + 1) % RN; ts_count++;
      ws_i = (ws_i + 1) % RN; ws_count++;
    }
    if (ph == piStep) {
      const long long qn = (long long)mcm::round_js((JS_PI - pr * dphi) / pitch);
      const int ip = (int)(((qn % N) + N) % N);
      const double dp = E.dR[ip] * 1e-3;
      if (lane == 0) {
        E.rings[4 * RN + wsp_i] = (float)(x - w * wnx - dp * vx);
        E.rings[5 * RN + wsp_i] = (float)(y - w * wny - dp * vy);
      }
      wsp_i = (wsp_i + 1) % RN; wsp_count++;
    }
    t += dt;
    // ---- step statistics (src/cutting.js:279-287)
    const double Fxy = mcm::hypot_js(Fx, Fy);
    st[S_N] += 1;
    st[S_SUMF] += Fxy; st[S_MAXF] = jmax(st[S_MAXF], Fxy);
    st[S_SUMFX] += Fx; st[S_SUMFY] += Fy; st[S_SUMFZ] += Fz; st[S_MAXFZ] = jmax(st[S_MAXFZ], -Fz);
    st[S_SUMT] += Tq; st[S_MAXT] = jmax(st[S_MAXT], Tq);
    st[S_SUMX] += x; st[S_SUMY] += y;
    st[S_MINX] = jmin(st[S_MINX], x); st[S_MAXX] = jmax(st[S_MAXX], x);
    st[S_MINY] = jmin(st[S_MINY], y); st[S_MAXY] = jmax(st[S_MAXY], y);
    if (Fsum > 0) { st[S_SUMZC] += Fr2z / Fsum * Fxy; st[S_WZC] += Fxy; }
    st[S_ENGT] += (double)__popcll(emask);
    if (++stepCount % decim == 0) {
      if (lane == 0) {
        E.hist[hist_i] = (float)Fx;
        E.hist[HN + hist_i] = (float)Fy;
        E.hist[2 * HN + hist_i] = (float)Fz;
        E.hist[3 * HN + hist_i] = (float)((x - w * wnx) * 1e6);
        E.hist[4 * HN + hist_i] = (float)((y - w * wny) * 1e6);
      }
      hist_i = (hist_i + 1) % HN; hist_count++;
    }
  }
  wave_sync();
  // ---- write the state back
  if (lane == 6) E.st[S_SUMHB] = chain;
  if (lane == 7) E.st[S_SUMB] = chain;
  if (lane == 8) E.st[S_SUMRB] = chain;
  if (isTooth) {
    E.toothVol[tooth] = tvol; E.toothPhi[tooth] = tphi; E.toothZ[tooth] = tz; E.toothEngaged[tooth] = tEng;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < S_COUNT; k++)
      if (k != S_SUMHB && k != S_SUMB && k != S_SUMRB) E.st[k] = st[k];
#pragma unroll
    for (int j = 0; j < MC_MAX_MODES; j++) { E.qx[j] = qx[j]; E.vxm[j] = vxm[j]; E.qy[j] = qy[j]; E.vym[j] = vym[j]; }
    E.x = E.x0 = x; E.y = E.y0 = y; E.vx = E.vx0 = vxs; E.vy = E.vy0 = vys;
    E.t = t; E.pr = pr; E.w = w; E.wv = wv;
    E.runMeanW = nSteps ? wSum / (double)nSteps : 0.0;
    E.stepCount = stepCount;
    E.hist_i = hist_i; E.hist_count = hist_count; E.ts_i = ts_i; E.ts_count = ts_count;
    E.ws_i = ws_i; E.ws_count = ws_count; E.wsp_i = wsp_i; E.wsp_count = wsp_count;
    E.nev = nev;
  }
}

// chatterIndex() over the device rings (src/cutting.js:299-318), one lane
__global__ __launch_bounds__(64) void engine_chatter_index(const EngDev* E, double* out) {
  if (threadIdx.x != 0) return;
  const int64_t n = E->ts_count < 24 ? E->ts_count : 24;
  if (n < 8) { out[0] = 0; out[1] = 0; out[2] = 0; return; }
  const float* tsx = E->rings;
  const float* tsy = E->rings + RN;
  double mx = 0, my = 0, v = 0;
  for (int64_t k = 0; k < n; k++) { const int64_t j = (E->ts_i - 1 - k + RN) % RN; mx += tsx[j]; my += tsy[j]; }
  mx /= n; my /= n;
  for (int64_t k = 0; k < n; k++) {
    const int64_t j = (E->ts_i - 1 - k + RN) % RN;
    const double dx = tsx[j] - mx, dy = tsy[j] - my;
    v += dx * dx + dy * dy;
  }
  const double once = __builtin_sqrt(v / n) * 1e6;
  const double r = mcm::round_js((double)n * E->spt / E->decim);
  const int64_t m = (double)E->hist_count < r ? E->hist_count : (int64_t)r;
  const float* hx = E->hist + 3 * HN;
  const float* hy = E->hist + 4 * HN;
  double ax = 0, ay = 0, w = 0;
  for (int64_t k = 0; k < m; k++) { const int64_t j = (E->hist_i - 1 - k + HN) % HN; ax += hx[j]; ay += hy[j]; }
  const double mm = m ? (double)m : 1.0;
  ax /= mm; ay /= mm;
  for (int64_t k = 0; k < m; k++) {
    const int64_t j = (E->hist_i - 1 - k + HN) % HN;
    const double dx = hx[j] - ax, dy = hy[j] - ay;
    w += dx * dx + dy * dy;
  }
  const double amp = __builtin_sqrt(w / mm);
  out[0] = once / (amp + 0.05);
  out[1] = amp;
  out[2] = once;
}

__global__ void fill_zero_f32(float* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

void new_stats(double* st) {
  for (int k = 0; k < S_COUNT; k++) st[k] = 0;
  st[S_MINX] = __builtin_inf(); st[S_MAXX] = -__builtin_inf();
  st[S_MINY] = __builtin_inf(); st[S_MAXY] = -__builtin_inf();
}

}  // namespace

struct mc_engine {
  EngDev h;                 // host mirror
  EngDev* d = nullptr;      // device copy
  size_t surf_bytes = 0;
  bool configured = false, engaged = false, wear = false;
  double wKte = 0, wKre = 0;
  mc_material mat{};
  int nonlinear = 1;
  int type = MC_TOOL_FLAT;
  std::vector<double> evh;  // chip events of the last run, read back on demand
  bool ev_fresh = false;
};

namespace {

int free_engine(mc_engine* e) {
  if (e->d) (void)hipFree(e->d);
  if (e->h.surf) (void)hipFree(e->h.surf);
  if (e->h.hist) (void)hipFree(e->h.hist);
  if (e->h.rings) (void)hipFree(e->h.rings);
  if (e->h.ev) (void)hipFree(e->h.ev);
  delete e;
  return MC_OK;
}

int zero_dev(float* p, int64_t n) {
  if (n <= 0) return MC_OK;
  const int blocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(fill_zero_f32, dim3(blocks), dim3(256), 0, mch::stream(), p, n);
  MCH_HIPCHK(hipGetLastError());
  return MC_OK;
}

// the run constants run() derives from the material, wear and engagement (src/cutting.js:146-165)
void run_constants(mc_engine* e) {
  EngDev& h = e->h;
  const bool nl = e->nonlinear != 0;
  h.Ktc = e->mat.Ktc; h.Krc = e->mat.Krc; h.Kac = e->mat.Kac; h.Kae = e->mat.Kae;
  h.Kte = e->wear ? e->wKte : e->mat.Kte;
  h.Kre = e->wear ? e->wKre : e->mat.Kre;
  h.mc = nl ? e->mat.mc : 0.0;
  h.pdK = nl ? e->mat.Ci * h.db * 1e-3 / (h.omega * 1e-3) : 0.0;
  const double TWO_PI = 2 * JS_PI;
  h.binScale = MC_BINS / TWO_PI;
  h.pitch = TWO_PI / h.N;
  h.dphi = h.omega * h.dt;
  h.hMax = jmax(0.1, 3 * h.fz);
}

int ensure_events(mc_engine* e, int64_t nsteps) {
  const int64_t cap = (int64_t)e->h.N * (nsteps / 2 + 2);
  if (cap <= e->h.ev_cap) return MC_OK;
  if (e->h.ev) { (void)hipFree(e->h.ev); e->h.ev = nullptr; e->h.ev_cap = 0; }
  if (hipMalloc(&e->h.ev, sizeof(double) * 3 * (size_t)cap) != hipSuccess)
    return mch::fail(MC_ERR_NOMEM, "chip event buffer (%lld events) failed", (long long)cap);
  e->h.ev_cap = cap;
  return MC_OK;
}

int check_engine(const mc_engine* e) {
  if (!e) return mch::fail(MC_ERR_ARG, "null engine");
  return MC_OK;
}

// the device pointer list of a launch (engines[i] -> its device struct)
struct PtrList {
  EngDev** d = nullptr;
  int cap = 0;
  int ensure(int n) {
    if (n <= cap) return MC_OK;
    if (d) (void)hipFree(d);
    d = nullptr; cap = 0;
    if (hipMalloc(&d, sizeof(EngDev*) * (size_t)n) != hipSuccess) return mch::fail(MC_ERR_NOMEM, "engine list");
    cap = n;
    return MC_OK;
  }
};
PtrList g_ptrs;

int upload(mc_engine* e) {
  MCH_HIPCHK(hipMemcpyAsync(e->d, &e->h, sizeof(EngDev), hipMemcpyHostToDevice, mch::stream()));
  return MC_OK;
}
int download(mc_engine* e) {
  MCH_HIPCHK(hipMemcpyAsync(&e->h, e->d, sizeof(EngDev), hipMemcpyDeviceToHost, mch::stream()));
  return MC_OK;
}

}  // namespace

extern "C" {

int mc_engine_create(mc_engine** out) {
  if (!out) return mch::fail(MC_ERR_ARG, "null output");
  *out = nullptr;
  if (int st = mch::init()) return st;
  std::lock_guard<std::mutex> lk(mch::lock());
  mc_engine* e = new mc_engine();
  std::memset(&e->h, 0, sizeof(EngDev));
  new_stats(e->h.st);
  e->h.wm = 1; e->h.wnx = 1;
  if (hipMalloc(&e->d, sizeof(EngDev)) != hipSuccess || hipMalloc(&e->h.hist, sizeof(float) * 5 * HN) != hipSuccess ||
      hipMalloc(&e->h.rings, sizeof(float) * 6 * RN) != hipSuccess) {
    free_engine(e);
    return mch::fail(MC_ERR_NOMEM, "engine allocation failed");
  }
  int st = zero_dev(e->h.hist, 5 * HN);
  if (!st) st = zero_dev(e->h.rings, 6 * RN);
  if (!st && hipStreamSynchronize(mch::stream()) != hipSuccess) st = mch::fail(MC_ERR_HIP, "engine set-up");
  if (st) { free_engine(e); return st; }
  *out = e;
  return MC_OK;
}

int mc_engine_destroy(mc_engine* e) {
  if (!e) return MC_OK;
  std::lock_guard<std::mutex> lk(mch::lock());
  (void)hipStreamSynchronize(mch::stream());
  return free_engine(e);
}

// configure(tool, modal, mat, S, fnMax), src/cutting.js:39-82
int mc_engine_configure(mc_engine* e, const mc_cell* c, double fnMax, mc_engine_info* info) {
  if (int st = check_engine(e)) return st;
  if (!c) return mch::fail(MC_ERR_ARG, "null cell");
  if (c->tool.N < 1 || c->tool.N > MC_MAX_FLUTES) return mch::fail(MC_ERR_LIMIT, "N=%d outside 1..%d", c->tool.N, MC_MAX_FLUTES);
  if (c->nmx < 1 || c->nmx > MC_MAX_MODES || c->nmy < 1 || c->nmy > MC_MAX_MODES)
    return mch::fail(MC_ERR_LIMIT, "modes per axis must be 1..%d", MC_MAX_MODES);
  if (!(c->tool.R > 0) || !(c->tool.fluteLen > 0) || !(c->S == c->S) || !(fnMax == fnMax))
    return mch::fail(MC_ERR_ARG, "invalid tool radius, flute length, speed or fnMax");
  EngDev& h = e->h;
  const double TWO_PI = 2 * JS_PI;
  h.N = c->tool.N; h.R = c->tool.R;
  h.tanB = mcm::tan_js(c->tool.helixDeg * JS_PI / 180);
  for (int i = 0; i < MC_MAX_FLUTES; i++)
    h.dR[i] = i < h.N ? c->tool.runout * mcm::cos_js(2 * JS_PI * i / h.N - c->tool.runoutPhase) : 0.0;
  h.S = jmax(100, c->S);
  h.omega = h.S * 2 * JS_PI / 60;
  h.Ttooth = 60 / (h.S * h.N);
  double fTop = -__builtin_inf();
  for (int i = 0; i < c->nmx; i++) fTop = jmax(fTop, c->mx[i].fn);
  for (int i = 0; i < c->nmy; i++) fTop = jmax(fTop, c->my[i].fn);
  const double sptd = jmax(48, std::ceil(h.Ttooth * jmax(fTop, fnMax) * 24));
  if (!(sptd * h.N <= (double)(1 << 24))) return mch::fail(MC_ERR_LIMIT, "%.0f steps per tooth too many", sptd);
  h.spt = (int)sptd;
  h.dt = h.Ttooth / h.spt;
  h.nmx = c->nmx; h.nmy = c->nmy;
  for (int i = 0; i < MC_MAX_MODES; i++) {
    h.kx[i] = h.mx[i] = h.cx[i] = h.ky[i] = h.my[i] = h.cy[i] = 0;
    h.qx[i] = h.vxm[i] = h.qy[i] = h.vym[i] = 0;
    h.mx[i] = h.my[i] = 1;
  }
  for (int i = 0; i < c->nmx; i++) {
    const double w = 2 * JS_PI * c->mx[i].fn;
    h.kx[i] = c->mx[i].k * 1e3; h.mx[i] = h.kx[i] / (w * w);
    h.cx[i] = 2 * c->mx[i].zeta * std::sqrt(h.kx[i] * h.mx[i]);
  }
  for (int i = 0; i < c->nmy; i++) {
    const double w = 2 * JS_PI * c->my[i].fn;
    h.ky[i] = c->my[i].k * 1e3; h.my[i] = h.ky[i] / (w * w);
    h.cy[i] = 2 * c->my[i].zeta * std::sqrt(h.ky[i] * h.my[i]);
  }
  h.x0 = h.y0 = h.vx0 = h.vy0 = 0;
  h.dz = jmax(0.25, c->tool.fluteLen / 48);
  const double ms = std::ceil(c->tool.fluteLen / h.dz);
  if (!(ms <= MS)) return mch::fail(MC_ERR_LIMIT, "%.0f slices (max %d)", ms, MS);
  h.maxSlices = (int)ms;
  h.spr = h.N * h.spt;
  const size_t sb = sizeof(float) * (size_t)h.maxSlices * h.spr;
  std::lock_guard<std::mutex> lk(mch::lock());
  if (sb > e->surf_bytes) {
    if (h.surf) (void)hipFree(h.surf);
    h.surf = nullptr; e->surf_bytes = 0;
    if (hipMalloc(&h.surf, sb) != hipSuccess) return mch::fail(MC_ERR_NOMEM, "delay line of %zu bytes failed", sb);
    e->surf_bytes = sb;
  }
  if (int st = zero_dev(h.surf, (int64_t)h.maxSlices * h.spr)) return st;
  h.pr = 0;
  h.x = h.y = h.vx = h.vy = h.t = 0;
  // setWork(null)
  h.runMeanW = 0; h.wk = 0; h.wm = 1; h.wc = 0; h.w = 0; h.wv = 0; h.wnx = 1; h.wny = 0;
  h.fza = 0;
  for (int i = 0; i < MC_MAX_FLUTES; i++) { h.toothEngaged[i] = 0; h.toothVol[i] = h.toothPhi[i] = h.toothZ[i] = 0; }
  h.decim = (int)jmax(1, mcm::round_js(h.spt / 40.0));
  const double rem = std::fmod(JS_PI, TWO_PI / h.N);
  h.piStep = (int)std::fmod(mcm::round_js((rem / (TWO_PI / h.N)) * h.spt), (double)h.spt);
  h.ws_i = h.ws_count = h.wsp_i = h.wsp_count = 0;
  h.stepCount = 0;
  e->mat = c->mat;
  e->nonlinear = c->nonlinear;
  e->type = c->tool.type;
  e->configured = true;
  e->engaged = false;
  if (info) {
    info->dt = h.dt; info->omega = h.omega; info->Ttooth = h.Ttooth; info->dz = h.dz; info->S = h.S;
    info->stepsPerTooth = h.spt; info->stepsPerRev = h.spr; info->decim = h.decim; info->piStep = h.piStep;
    info->maxSlices = h.maxSlices; info->nmx = h.nmx; info->nmy = h.nmy; info->pad = 0;
  }
  MCH_HIPCHK(hipStreamSynchronize(mch::stream()));
  return MC_OK;
}

int mc_engine_set_wear(mc_engine* e, double Kte, double Kre) {
  if (int st = check_engine(e)) return st;
  e->wear = true; e->wKte = Kte; e->wKre = Kre;
  return MC_OK;
}

// setWork(wall), src/cutting.js:88-96
int mc_engine_set_work(mc_engine* e, const mc_wall* w) {
  if (int st = check_engine(e)) return st;
  EngDev& h = e->h;
  if (!w) { h.runMeanW = 0; h.wk = 0; h.wm = 1; h.wc = 0; h.w = 0; h.wv = 0; h.wnx = 1; h.wny = 0; return MC_OK; }
  if (truthy(h.wk) && w->nx * h.wnx + w->ny * h.wny < 0.8) { h.w = 0; h.wv = 0; }
  h.wk = w->k * 1e3; h.wm = w->m; h.wc = 2 * w->zeta * std::sqrt(h.wk * w->m);
  h.wnx = w->nx; h.wny = w->ny;
  return MC_OK;
}

// setEngagement(u, fz, tipZ, ap), src/cutting.js:98-124
int mc_engine_set_engagement(mc_engine* e, double ux, double uy, double fz, double tipZ, double ap,
                             const double* sliceR, const double* sliceKap, int32_t* nz_out) {
  if (int st = check_engine(e)) return st;
  if (!e->configured) return mch::fail(MC_ERR_ARG, "setEngagement before configure");
  if ((sliceR == nullptr) != (sliceKap == nullptr)) return mch::fail(MC_ERR_ARG, "sliceR and sliceKap go together");
  EngDev& h = e->h;
  h.ux = ux; h.uy = uy; h.ex = -uy; h.ey = ux;
  h.fz = fz; h.tipZ = tipZ;
  const int nz = ap <= 0 ? 0 : (int)jmin(h.maxSlices, std::ceil(ap / h.dz));
  h.nz = nz;
  h.db = h.dz;
  for (int k = 0; k < nz; k++) {
    h.sB[k] = jmin(h.dz, ap - k * h.dz);
    const double zr = k * h.dz + 0.5 * h.sB[k];
    h.sZ[k] = tipZ + zr;
    double r = h.R, kap = 1;
    if (sliceR) { r = sliceR[k]; kap = sliceKap[k]; }
    else if (e->type == MC_TOOL_BALL && zr < h.R) {
      const double q = h.R - zr;
      r = std::sqrt(jmax(1e-6, h.R * h.R - q * q));
      kap = r / h.R;
    }
    h.sR[k] = r; h.sKap[k] = kap;
    h.sLag[k] = zr * h.tanB / h.R;
  }
  e->engaged = true;
  if (nz_out) *nz_out = nz;
  return MC_OK;
}

// setEndEngagement(fza, rOut, zAt), src/cutting.js:130-143
int mc_engine_set_end_engagement(mc_engine* e, double fza, double rOut, const double* rings) {
  if (int st = check_engine(e)) return st;
  EngDev& h = e->h;
  h.fza = fza > 0 ? fza : 0;
  h.endR = rOut;
  const double dr = rOut / ERINGS;
  for (int j = 0; j < ERINGS; j++) {
    if (rings) {
      h.ringR[j] = rings[j]; h.ringCos[j] = rings[ERINGS + j];
      h.ringZ[j] = rings[2 * ERINGS + j]; h.ringDs[j] = rings[3 * ERINGS + j];
    } else {
      const double r = (j + 0.5) * dr;
      const double slope = (0.0 - 0.0) / (0.5 * dr);
      const double c = jmax(0.22, 1 / std::sqrt(1 + slope * slope));
      h.ringR[j] = r; h.ringCos[j] = c; h.ringDs[j] = dr / c; h.ringZ[j] = 0;
    }
  }
  return MC_OK;
}

int mc_engine_set_maps(mc_engine* e, const float* Hb, const float* Eb) {
  if (int st = check_engine(e)) return st;
  if (Hb) std::memcpy(e->h.Hb, Hb, sizeof(e->h.Hb));
  if (Eb) std::memcpy(e->h.Eb, Eb, sizeof(e->h.Eb));
  return MC_OK;
}

int mc_engine_get_state(const mc_engine* e, mc_engine_state* s) {
  if (!e || !s) return mch::fail(MC_ERR_ARG, "null engine or state");
  const EngDev& h = e->h;
  std::memset(s, 0, sizeof(*s));
  s->x = h.x; s->y = h.y; s->vx = h.vx; s->vy = h.vy; s->w = h.w; s->wv = h.wv; s->t = h.t;
  s->runMeanW = h.runMeanW; s->wk = h.wk; s->wnx = h.wnx; s->wny = h.wny;
  s->stepCount = h.stepCount; s->pr = h.pr; s->decim = h.decim; s->nz = h.nz;
  s->hist_i = h.hist_i; s->hist_count = h.hist_count; s->ts_i = h.ts_i; s->ts_count = h.ts_count;
  s->ws_i = h.ws_i; s->ws_count = h.ws_count; s->wsp_i = h.wsp_i; s->wsp_count = h.wsp_count;
  return MC_OK;
}

int mc_engine_set_state(mc_engine* e, const mc_engine_state* s) {
  if (!e || !s) return mch::fail(MC_ERR_ARG, "null engine or state");
  if (s->decim < 1) return mch::fail(MC_ERR_ARG, "decim must be >= 1");
  if (s->hist_i < 0 || s->hist_i >= HN) return mch::fail(MC_ERR_ARG, "hist index outside 0..%d", HN - 1);
  EngDev& h = e->h;
  h.x = s->x; h.y = s->y; h.vx = s->vx; h.vy = s->vy; h.w = s->w; h.wv = s->wv; h.t = s->t;
  h.decim = s->decim; h.hist_i = s->hist_i;
  return MC_OK;
}

// run(nSteps), src/cutting.js:146-296, for n engines in one launch
int mc_engine_run_batch(mc_engine* const* es, const int64_t* nSteps, int n) {
  if (n < 0 || (n > 0 && (!es || !nSteps))) return mch::fail(MC_ERR_ARG, "null engines or negative count");
  if (n == 0) return MC_OK;
  for (int i = 0; i < n; i++) {
    mc_engine* e = es[i];
    if (!e) return mch::fail(MC_ERR_ARG, "engine %d is null", i);
    if (!e->configured || !e->engaged) return mch::fail(MC_ERR_ARG, "engine %d: run needs configure and then setEngagement", i);
    if (nSteps[i] < 0) return mch::fail(MC_ERR_ARG, "engine %d: negative step count", i);
    if (e->h.N * (e->h.nz + ERINGS) > SLOT_CAP) return mch::fail(MC_ERR_LIMIT, "engine %d: too many visits per step", i);
    for (int j = 0; j < i; j++)
      if (es[j] == e) return mch::fail(MC_ERR_ARG, "engine %d appears twice in the batch", i);
  }
  std::lock_guard<std::mutex> lk(mch::lock());
  if (int st = g_ptrs.ensure(n)) return st;
  std::vector<EngDev*> ptrs((size_t)n);
  for (int i = 0; i < n; i++) {
    mc_engine* e = es[i];
    if (int st = ensure_events(e, nSteps[i])) return st;
    run_constants(e);
    e->h.nsteps = nSteps[i];
    e->h.nev = 0;
    if (int st = upload(e)) return st;
    ptrs[(size_t)i] = e->d;
  }
  MCH_HIPCHK(hipMemcpyAsync(g_ptrs.d, ptrs.data(), sizeof(EngDev*) * (size_t)n, hipMemcpyHostToDevice, mch::stream()));
  hipLaunchKernelGGL(engine_run, dim3(n), dim3(64), 0, mch::stream(), (EngDev* const*)g_ptrs.d);
  MCH_HIPCHK(hipGetLastError());
  for (int i = 0; i < n; i++)
    if (int st = download(es[i])) return st;
  MCH_HIPCHK(hipStreamSynchronize(mch::stream()));
  for (int i = 0; i < n; i++) {
    es[i]->ev_fresh = false;
    if (es[i]->h.nev > es[i]->h.ev_cap) return mch::fail(MC_ERR_LIMIT, "engine %d: chip events overflowed", i);
  }
  return MC_OK;
}

int mc_engine_run(mc_engine* e, int64_t nSteps) {
  mc_engine* one[1] = {e};
  return mc_engine_run_batch(one, &nSteps, 1);
}

int mc_engine_read_stats(mc_engine* e, mc_engine_stats* out, int take) {
  if (int st = check_engine(e)) return st;
  if (out) std::memcpy(out, e->h.st, sizeof(mc_engine_stats));
  if (take) new_stats(e->h.st);
  return MC_OK;
}

int mc_engine_chip_events(const mc_engine* e, mc_chip_event* out, int64_t cap, int64_t* n) {
  if (!e) return mch::fail(MC_ERR_ARG, "null engine");
  const int64_t m = e->h.nev;
  if (n) *n = m;
  const int64_t k = std::min<int64_t>(m, std::max<int64_t>(cap, 0));
  if (k <= 0 || !out) return MC_OK;
  mc_engine* me = const_cast<mc_engine*>(e);
  std::lock_guard<std::mutex> lk(mch::lock());
  if (!me->ev_fresh) {
    me->evh.resize((size_t)(3 * m));
    MCH_HIPCHK(hipMemcpy(me->evh.data(), e->h.ev, sizeof(double) * 3 * (size_t)m, hipMemcpyDeviceToHost));
    me->ev_fresh = true;
  }
  const EngDev& h = e->h;
  for (int64_t i = 0; i < k; i++) {
    out[i].vol = me->evh[(size_t)(3 * i)]; out[i].phi = me->evh[(size_t)(3 * i + 1)];
    out[i].z = me->evh[(size_t)(3 * i + 2)];
    out[i].ux = h.ux; out[i].uy = h.uy; out[i].vx = h.ex; out[i].vy = h.ey;
  }
  return MC_OK;
}

int mc_engine_chatter_index(mc_engine* e, double out[3]) {
  if (int st = check_engine(e)) return st;
  if (!out) return mch::fail(MC_ERR_ARG, "null output");
  std::lock_guard<std::mutex> lk(mch::lock());
  double* d = nullptr;
  if (hipMalloc(&d, 3 * sizeof(double)) != hipSuccess) return mch::fail(MC_ERR_NOMEM, "chatter index buffer");
  int st = upload(e);
  if (!st) {
    hipLaunchKernelGGL(engine_chatter_index, dim3(1), dim3(64), 0, mch::stream(), (const EngDev*)e->d, d);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, d, 3 * sizeof(double), hipMemcpyDeviceToHost, mch::stream()) != hipSuccess ||
        hipStreamSynchronize(mch::stream()) != hipSuccess)
      st = mch::fail(MC_ERR_HIP, "chatter index kernel failed");
  }
  (void)hipFree(d);
  return st;
}

int mc_engine_read_rings(const mc_engine* e, float* hist, float* ts, float* ws, float* wsp) {
  if (!e) return mch::fail(MC_ERR_ARG, "null engine");
  std::lock_guard<std::mutex> lk(mch::lock());
  std::vector<float> r(6 * RN);
  if (hist) MCH_HIPCHK(hipMemcpy(hist, e->h.hist, sizeof(float) * 5 * HN, hipMemcpyDeviceToHost));
  if (ts || ws || wsp) {
    MCH_HIPCHK(hipMemcpy(r.data(), e->h.rings, sizeof(float) * 6 * RN, hipMemcpyDeviceToHost));
    if (ts) std::memcpy(ts, r.data(), sizeof(float) * 2 * RN);
    if (ws) std::memcpy(ws, r.data() + 2 * RN, sizeof(float) * 2 * RN);
    if (wsp) std::memcpy(wsp, r.data() + 4 * RN, sizeof(float) * 2 * RN);
  }
  return MC_OK;
}

}  // extern "C"
