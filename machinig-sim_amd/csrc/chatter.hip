// chatter.hip — persistent time-domain chatter search for gfx950.
//
// One GROUP of G lanes (G = 8..64, a fraction of a 64-lane wavefront) owns one stability
// cell at a time and runs its whole search: the chatterAt trials of timeDomainLimit
// (src/stabtd.js:14-32), each a fresh CuttingEngine run (src/cutting.js:146-296) of
// spt*N*revs steps, followed by the chatter index (src/cutting.js:299-318) and the
// bisection update.  Groups pull cells from a device work queue, so cells of uneven
// length (revolutions, flutes, modes, depths) share one launch and one wave never waits
// for another's cell.
//
// Per step, the lanes of a group split the N*nz (tooth, slice) visits; each lane computes
// its visits' engagement, chip thickness and force, writes the delay line and leaves its
// force terms in LDS.  Every lane then sums the terms serially in the reference's
// (tooth, slice) order — the FP64 result is bit-identical to the JavaScript loop — and
// advances the multi-mode modal state redundantly, so the state stays wave-resident
// without any cross-lane broadcast.  The delay line (nz * stepsPerRev float32 values,
// up to 225 KB) lives in global memory, one region per group; the values a step needs
// were written one tooth period earlier, so they are prefetched a step ahead.
#include <hip/hip_runtime.h>

#include "mc_dev.h"
#include "mc_math.h"

using namespace mcd;

namespace {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double js_min(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double js_max(double a, double b) { return a > b ? a : b; }

struct GroupLds {
  double2* term;   // [nv_cap] force terms of this step's visits
  double* sLag;    // [nz_cap] helix lag of each slice
  double* sZ;      // [nz_cap] slice height
  double* sB;      // [nz_cap] slice width
  double* sKap;    // [nz_cap] edge inclination factor
  double* sR;      // [nz_cap] slice radius
  double* modal;   // kx mx cx ky my cy, MC_MAX_MODES each
  double* dR;      // [MC_MAX_FLUTES] tooth radial offsets
  double* bcd;     // 4 doubles of broadcast
  float* Hb;       // [MC_BINS]
  float* tsx;      // [TS]
  float* tsy;      // [TS]
  unsigned long long* eng;
  int* bci;        // 4 ints of broadcast
};

__device__ __forceinline__ GroupLds carve(unsigned char* base, int nv_cap, int nz_cap) {
  GroupLds L;
  L.term = reinterpret_cast<double2*>(base);
  double* d = reinterpret_cast<double*>(base + 16 * (size_t)nv_cap);
  L.sLag = d; d += nz_cap;
  L.sZ = d; d += nz_cap;
  L.sB = d; d += nz_cap;
  L.sKap = d; d += nz_cap;
  L.sR = d; d += nz_cap;
  L.modal = d; d += 6 * MC_MAX_MODES;
  L.dR = d; d += MC_MAX_FLUTES;
  L.bcd = d; d += 4;
  L.eng = reinterpret_cast<unsigned long long*>(d); d += 1;
  L.bci = reinterpret_cast<int*>(d);
  float* f = reinterpret_cast<float*>(L.bci + 4);
  L.Hb = f; f += MC_BINS;
  L.tsx = f; f += TS;
  L.tsy = f;
  return L;
}

}  // namespace

// bytes of LDS a group needs (kept in sync with carve())
int mcd_group_lds_bytes(int nv_cap, int nz_cap) {
  size_t b = 16 * (size_t)nv_cap + 8 * (5 * (size_t)nz_cap + 6 * MC_MAX_MODES + MC_MAX_FLUTES + 4 + 1) +
             4 * 4 + 4 * (MC_BINS + 2 * TS);
  return (int)((b + 15) & ~(size_t)15);
}

template <int G, int MAXP>
__global__ __launch_bounds__(64) void chatter_search(LaunchDev P) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  constexpr int GPW = 64 / G;
  const int lane = threadIdx.x;
  const int g = lane % G;
  const int grp = lane / G;
  const int64_t gid = (int64_t)blockIdx.x * GPW + grp;
  GroupLds L = carve(lds + grp * P.grp_bytes, P.nv_cap, P.nz_cap);
  float* ring = P.ring + gid * P.ring_cap;
  float2* hist = P.hist + gid * HIST;

  // ---- task (bisection) state, uniform within the group
  int task = -1, mode = 0, phase = 0, iters = 0, it = 0, fell = 0, trials = 0;
  bool alive = true;
  double lo = 0, hi = 0, aMin = 0, aMax = 0, tol = 0, ap = 0, lo_hint = 0, hi_hint = 0;
  long long steps_acc = 0, visits_acc = 0, my_eng = 0;
  // ---- trial state
  const CellDev* C = nullptr;
  int N = 0, nz = 0, NV = 0, npass = 0, spt = 0, spr = 0, decim = 0, nmx = 0, nmy = 0;
  long long sLeft = 0, sTotal = 0;
  int pr = 0, ph = 0, dc = 0, ti = 0, tcount = 0, hpos = 0, hcount = 0;
  double x = 0, y = 0, vxs = 0, vys = 0;
  double qx[MC_MAX_MODES], vxm[MC_MAX_MODES], qy[MC_MAX_MODES], vym[MC_MAX_MODES];
  double fz = 0, dphi = 0, pitch = 0, binScale = 0, hMax = 0, dt = 0, db = 0;
  double Ktc = 0, Krc = 0, Kte = 0, Kre = 0, mcx = 0, pdK = 0;
  float cur[MAXP];
  // feed direction: chatterAt's u = [1, 0], or the task's; v = [-u1, u0] (src/stabtd.js:18,
  // cutting.js:99)
  double ux = 1.0, uy = 0.0, vx = -0.0, vy = 1.0;
  bool simeng = false;         // the task is a simulator sub-step's engine
  double tipZs = 0, kvx = 0.01, kvy = 0.0;
  const double TWO_PI = mcm::TWO_PI;

  while (true) {
    if (alive && sLeft == 0) {
      // ================= trial end, bisection update, next trial (rare, divergent)
      if (task >= 0) {
        if (g == 0) {
          // chatterIndex(), src/cutting.js:299-318
          double index = 0, amp = 0, once = 0;
          const int n = tcount < 24 ? tcount : 24;
          if (n >= 8) {
            double mx = 0, my = 0, v = 0;
            for (int k = 0; k < n; k++) { const int j = (ti - 1 - k) & (TS - 1); mx += L.tsx[j]; my += L.tsy[j]; }
            mx /= n; my /= n;
            for (int k = 0; k < n; k++) {
              const int j = (ti - 1 - k) & (TS - 1);
              const double dx = L.tsx[j] - mx, dy = L.tsy[j] - my;
              v += dx * dx + dy * dy;
            }
            once = __builtin_sqrt(v / n) * 1e6;
            const double r = mcm::round_js((double)n * (double)spt / (double)decim);
            const int m = (double)hcount < r ? hcount : (int)r;
            double ax = 0, ay = 0, w = 0;
            for (int k = 0; k < m; k++) { const float2 h = hist[(hpos - 1 - k) & (HIST - 1)]; ax += h.x; ay += h.y; }
            const double mm = m ? (double)m : 1.0;
            ax /= mm; ay /= mm;
            for (int k = 0; k < m; k++) {
              const float2 h = hist[(hpos - 1 - k) & (HIST - 1)];
              const double dx = h.x - ax, dy = h.y - ay;
              w += dx * dx + dy * dy;
            }
            amp = __builtin_sqrt(w / mm);
            index = once / (amp + 0.05);
          }
          L.bcd[0] = index; L.bcd[1] = amp; L.bcd[2] = once;
        }
        wave_sync();
        const double index = L.bcd[0];
        const bool chat = index > MC_CHATTER_IDX;
        trials++;
        steps_acc += sTotal;
        visits_acc += (long long)NV * sTotal;
        bool done = false;
        if (mode == TASK_SINGLE) {
          done = true;
        } else if (mode == TASK_BISECT) {
          if (phase == 0) {
            if (!chat) { lo = aMax; hi = __builtin_inf(); done = true; }
            else { lo = aMin; hi = aMax; it = 0; phase = 1; done = iters <= 0; }
          } else {
            if (chat) hi = ap; else lo = ap;
            done = ++it >= iters;
          }
          if (!done) ap = __builtin_sqrt(lo * hi);
        } else {  // TASK_BRACKET
          if (phase == 10) {
            if (chat) { mode = TASK_BISECT; phase = 0; fell = 1; ap = aMax; }
            else { phase = 11; ap = hi_hint; }
          } else if (phase == 11) {
            if (!chat) { mode = TASK_BISECT; phase = 0; fell = 1; ap = aMax; }
            else {
              lo = lo_hint; hi = hi_hint; it = 0; phase = 12;
              if (it < iters && hi - lo > tol) ap = __builtin_sqrt(lo * hi); else done = true;
            }
          } else {
            if (chat) hi = ap; else lo = ap;
            ++it;
            if (it < iters && hi - lo > tol) ap = __builtin_sqrt(lo * hi); else done = true;
          }
        }
        if (done) {
          atomicAdd(L.eng, (unsigned long long)my_eng);
          wave_sync();
          if (g == 0) {
            ResultDev& R = P.out[task];
            const bool open = mode == TASK_BISECT && hi == __builtin_inf();
            R.open = open;
            R.lo = lo; R.hi = hi;
            R.alim = mode == TASK_SINGLE ? 0.0 : (open ? aMax : __builtin_sqrt(lo * hi));
            R.index = index; R.amp = L.bcd[1]; R.once = L.bcd[2];
            R.chatter = chat; R.nz = nz; R.fell_back = fell; R.trials = trials;
            R.steps = steps_acc; R.visits = visits_acc; R.engaged = (long long)*L.eng;
          }
          wave_sync();
          task = -1;
        }
      }
      if (task < 0) {
        // ---- next task from the queue
        if (g == 0) L.bci[0] = atomicAdd(P.queue, 1);
        wave_sync();
        task = L.bci[0];
        if (task >= P.ntasks) {
          alive = false;
        } else {
          const TaskDev T = P.tasks[task];
          C = P.cells + T.cell;
          mode = T.mode; iters = T.iters; aMin = T.aMin; aMax = T.aMax; tol = T.tol;
          lo_hint = T.lo; hi_hint = T.hi;
          it = 0; fell = 0; trials = 0; steps_acc = 0; visits_acc = 0; my_eng = 0;
          simeng = mode == TASK_SINGLE && T.sim != 0;
          ux = simeng ? T.ux : 1.0; uy = simeng ? T.uy : 0.0;
          vx = -uy; vy = ux;
          tipZs = T.tipZ;
          kvx = simeng ? T.vx0 : 0.01; kvy = simeng ? T.vy0 : 0.0;
          if (mode == TASK_SINGLE) ap = T.ap;
          else if (mode == TASK_BISECT) { phase = 0; ap = aMax; }
          else { phase = 10; ap = lo_hint; }
          if (g == 0) *L.eng = 0;
          N = C->N; spt = C->spt; spr = C->spr; decim = C->decim; nmx = C->nmx; nmy = C->nmy;
          fz = C->fz; dphi = C->dphi; pitch = C->pitch; binScale = C->binScale; hMax = C->hMax;
          dt = C->dt; Ktc = C->Ktc; Krc = C->Krc; Kte = C->Kte; Kre = C->Kre; mcx = C->mc;
          for (int j = g; j < 6 * MC_MAX_MODES; j += G) {
            const int a = j / MC_MAX_MODES, b = j % MC_MAX_MODES;
            const double* src = a == 0 ? C->kx : a == 1 ? C->mx : a == 2 ? C->cx
                              : a == 3 ? C->ky : a == 4 ? C->my : C->cy;
            L.modal[j] = src[b];
          }
          for (int j = g; j < MC_MAX_FLUTES; j += G) L.dR[j] = C->dR[j];
          for (int j = g; j < MC_BINS; j += G) L.Hb[j] = C->Hb[j];
        }
      }
      if (alive) {
        // ---- fresh engine at depth ap: configure + setEngagement (src/stabtd.js:15-19)
        // chatterAt lengthens the flutes to ap + 1 (src/stabtd.js:16); the simulator's
        // engine keeps the tool's own (src/sim.js:255)
        const double fl = simeng ? C->toolFluteLen : js_max(C->toolFluteLen, ap + 1);
        const double tz = simeng ? tipZs : -ap;
        const double dz = js_max(0.25, fl / 48);
        const int maxSlices = (int)__builtin_ceil(fl / dz);
        nz = ap <= 0 ? 0 : (int)js_min((double)maxSlices, __builtin_ceil(ap / dz));
        if (nz > P.nz_cap) nz = P.nz_cap;  // host sizes nz_cap from the largest depth
        NV = N * nz;
        npass = (NV + G - 1) / G;
        db = dz;
        pdK = C->Ci * db * 1e-3 / (C->omega * 1e-3);
        const double R = C->R;
        for (int k = g; k < nz; k += G) {
          const double b = js_min(dz, ap - k * dz);
          const double zr = k * dz + 0.5 * b;
          double r = R, kap = 1;
          if (C->ball && zr < R) {
            const double t = R - zr;
            r = __builtin_sqrt(js_max(1e-6, R * R - t * t));
            kap = r / R;
          }
          L.sB[k] = b; L.sZ[k] = tz + zr; L.sR[k] = r; L.sKap[k] = kap;
          L.sLag[k] = zr * C->tanB / R;
        }
        const int rn = nz * spr;
        for (int j = g; j < rn; j += G) ring[j] = 0.0f;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < MC_MAX_MODES; j++) { qx[j] = 0; vxm[j] = 0; qy[j] = 0; vym[j] = 0; }
        // a velocity set from outside (chatterAt's kick vx = 0.01) enters the first mode
        // of each axis (src/stabtd.js:19, cutting.js:164)
        vxm[0] = 0.0 + (kvx - 0.0);
        vym[0] = 0.0 + (kvy - 0.0);
        x = 0; y = 0; vxs = kvx; vys = kvy;
        pr = 0; ph = 0; dc = 0; ti = 0; tcount = 0; hpos = 0; hcount = 0;
        sTotal = sLeft = (long long)spt * N * C->revs;
#pragma unroll
        for (int p = 0; p < MAXP; p++) cur[p] = 0.0f;
      }
    }
    if (!__any(alive)) break;
    if (!alive || sLeft == 0) continue;

    // ===================== one dynamics step (src/cutting.js:168-293, side edges only)
    const double theta = (double)pr * dphi;
    const double xm = x * 1e3, ym = y * 1e3;   // rigid workpiece: rx = x - 0*1 = x
    const double rvx = vxs, rvy = vys;
    const int prn = pr + 1 == spr ? 0 : pr + 1;
    float nxt[MAXP];
#pragma unroll
    for (int p = 0; p < MAXP; p++) {
      nxt[p] = 0.0f;
      const int v = g + p * G;
      if (p < npass && v < NV) {
        const int i = v / nz, k = v - i * nz;
        int pi = prn + i * spt;
        if (pi >= spr) pi -= spr;
        nxt[p] = ring[pi * nz + k];
      }
    }
#pragma unroll
    for (int p = 0; p < MAXP; p++) {
      const int v = g + p * G;
      if (p < npass && v < NV) {
        const int i = v / nz, k = v - i * nz;
        double phi = mcm::fmod_2pi((theta + (double)i * pitch) - L.sLag[k]);
        if (phi < 0) phi += TWO_PI;
        double sn, cs;
        mcm::sincos_js(phi, sn, cs);
        const double nxg = sn * ux + cs * vx, nyg = sn * uy + cs * vy;
        const double rn = xm * nxg + ym * nyg + L.dR[i];
        int pi = pr + i * spt;
        if (pi >= spr) pi -= spr;
        const double prev = (double)cur[p];
        const double feedH = fz * sn * L.sKap[k];
        const int bin = (int)(phi * binScale);
        const bool below = bin < MC_BINS ? ((double)L.Hb[bin] <= L.sZ[k]) : false;
        double tX = 0, tY = 0;
        float store;
        if (below || sn <= 0) {
          store = (float)rn;
        } else {
          double h = feedH + rn - prev;
          if (h <= 0) {
            store = (float)(prev - feedH);
          } else {
            store = (float)rn;
            my_eng++;
            if (h > hMax) h = hMax;
            const double se = (mcx != 0 && mcx == mcx) ? mcm::pow_js(0.05 / (h > 0.002 ? h : 0.002), mcx) : 1.0;
            const double bk = L.sB[k];
            const double Ft = Ktc * se * h * bk + Kte * bk;
            double Fr = Krc * se * h * bk + Kre * bk;
            if (pdK != 0 && pdK == pdK) Fr += pdK * (bk / db) / L.sR[k] * (rvx * nxg + rvy * nyg);
            const double txg = cs * ux - sn * vx, tyg = cs * uy - sn * vy;
            tX = -Ft * txg - Fr * nxg;
            tY = -Ft * tyg - Fr * nyg;
          }
        }
        ring[pi * nz + k] = store;
        L.term[v] = make_double2(tX, tY);
      }
    }
    wave_sync();
    // serial sum in (tooth, slice) order, as Fx += ... in the reference loop
    double Fx = 0, Fy = 0;
    for (int v = 0; v < NV; v++) {
      const double2 t = L.term[v];
      Fx += t.x;
      Fy += t.y;
    }
    wave_sync();
    // modal integration, semi-implicit Euler (src/cutting.js:247-255)
    x = 0; y = 0; vxs = 0; vys = 0;
#pragma unroll
    for (int j = 0; j < MC_MAX_MODES; j++) {
      if (j < nmx) {
        const double k = L.modal[j], m = L.modal[MC_MAX_MODES + j], c = L.modal[2 * MC_MAX_MODES + j];
        vxm[j] += (Fx - c * vxm[j] - k * qx[j]) / m * dt;
        qx[j] += vxm[j] * dt;
        x += qx[j];
        vxs += vxm[j];
      }
    }
#pragma unroll
    for (int j = 0; j < MC_MAX_MODES; j++) {
      if (j < nmy) {
        const double k = L.modal[3 * MC_MAX_MODES + j], m = L.modal[4 * MC_MAX_MODES + j],
                     c = L.modal[5 * MC_MAX_MODES + j];
        vym[j] += (Fy - c * vym[j] - k * qy[j]) / m * dt;
        qy[j] += vym[j] * dt;
        y += qy[j];
        vys += vym[j];
      }
    }
    // samples for the chatter index (src/cutting.js:265-268, 288-292)
    pr = prn;
    ph = ph + 1 == spt ? 0 : ph + 1;
    if (ph == 0) {
      if (g == 0) { L.tsx[ti] = (float)x; L.tsy[ti] = (float)y; }
      ti = (ti + 1) & (TS - 1);
      tcount++;
    }
    dc = dc + 1 == decim ? 0 : dc + 1;
    if (dc == 0) {
      if (g == 0) hist[hpos] = make_float2((float)(x * 1e6), (float)(y * 1e6));
      hpos = (hpos + 1) & (HIST - 1);
      hcount++;
    }
    if (P.trace && g == 0) {
      const long long s = sTotal - sLeft;
      if (s < P.trace_steps) {
        double* t = P.trace + 4 * s;
        t[0] = Fx; t[1] = Fy; t[2] = x; t[3] = y;
      }
    }
#pragma unroll
    for (int p = 0; p < MAXP; p++) cur[p] = nxt[p];
    sLeft--;
  }
}

// launch with G lanes per cell; passes of G lanes cover up to 8*G visits per step
int mcd_launch_chatter(const LaunchDev& P, int G, int blocks, hipStream_t stream) {
  const size_t lds = (size_t)(64 / G) * P.grp_bytes;
  switch (G) {
    case 8: hipLaunchKernelGGL((chatter_search<8, 8>), dim3(blocks), dim3(64), lds, stream, P); break;
    case 16: hipLaunchKernelGGL((chatter_search<16, 8>), dim3(blocks), dim3(64), lds, stream, P); break;
    case 32: hipLaunchKernelGGL((chatter_search<32, 8>), dim3(blocks), dim3(64), lds, stream, P); break;
    case 64: hipLaunchKernelGGL((chatter_search<64, 8>), dim3(blocks), dim3(64), lds, stream, P); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
