// search.hip — speculative, wave-per-cell time-domain stability search for gfx950.
//
// timeDomainLimit (src/stabtd.js:24-32) checks the depth aMax, then bisects
// geometrically: every trial is a full chatterAt simulation (src/cutting.js:146-318) of
// spt*N*revs steps whose next depth depends on the previous verdict.  Here one wavefront
// owns one cell and, per ROUND, evaluates the next L levels of the bisection tree at once
// (1, 3 or 7 trials, one per lane SLOT of 64, 16 or 8 lanes): it then walks the tree with
// the verdicts, which reproduces the serial bisection exactly (same depths, same
// brackets), in ceil(levels / L) rounds instead of one trial per level.  Every trial of a
// round has the same length (same cell), so the slots advance in lockstep, and all cell
// constants are wave-uniform.  The aMax check is a separate task of its own wave, run
// concurrently; the host combines the two (stable at aMax -> open result).
//
// The slot size is chosen per round from the trial depths so that the N*nz (tooth,
// slice) visits of every slot fit in at most MAXP passes of its lanes.  Per step the
// lanes of a slot split the visits (engagement, chip thickness, forces, delay line),
// leave their force terms in LDS, then every lane sums its slot's terms serially in the
// reference's (tooth, slice) order (bit-identical FP64) and advances that slot's modal
// state.  The delay-line values a step reads were written a tooth period earlier and are
// prefetched a step ahead; the edge angles, sines and cosines of step s+1 (independent of
// the vibration) are computed while step s's forces are evaluated.
#include <hip/hip_runtime.h>

#include "mc_dev.h"
#include "mc_math.h"

using namespace mcd;

namespace {

constexpr int MAXP = 2;        // passes of a slot's lanes per step
constexpr int MAXSLOT = 8;     // slots of a wave (8 lanes each at most 8 slots)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double unid(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double js_min(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double js_max(double a, double b) { return a > b ? a : b; }

// slice count of chatterAt's engine at depth ap (src/stabtd.js:16, cutting.js:63-64,101)
__device__ __forceinline__ int slices_at(double toolFL, double ap, double& dz) {
  const double fl = js_max(toolFL, ap + 1);
  dz = js_max(0.25, fl / 48);
  const int maxSlices = (int)__builtin_ceil(fl / dz);
  return ap <= 0 ? 0 : (int)js_min((double)maxSlices, __builtin_ceil(ap / dz));
}

struct SlotLds {
  double2* term;   // [nv_cap]
  double* sLag;    // [nz_cap]
  double* sZ;
  double* sB;
  double* sKap;
  double* sR;
  float* tsx;      // [TS]
  float* tsy;
};

struct WaveLds {
  double2 geo[MAXP][64];            // next step's (sin, cos) of each lane's visits
  int gbin[MAXP][64];               // next step's engagement-map bin
  double mq[MAXSLOT][2 * MC_MAX_MODES], mv[MAXSLOT][2 * MC_MAX_MODES];  // modal q, v per slot
  double node_ap[MAXSLOT], node_lo[MAXSLOT], node_hi[MAXSLOT];
  double idx[MAXSLOT], amp[MAXSLOT], once[MAXSLOT];
  long long eng[MAXSLOT];
  int nz[MAXSLOT];
  float Hb[MC_BINS];
  int task;
};

}  // namespace

// LDS of a wave: the per-wave block, 64*MAXP force terms (a slot of SG lanes owns
// SG*MAXP of them), then per slot nz_cap slice records and its tooth-sample ring
int mcd_search_lds_bytes(int nv_cap, int nz_cap) {
  (void)nv_cap;
  const size_t wave = (sizeof(WaveLds) + 15) & ~(size_t)15;
  const size_t slot = (8 * 5 * (size_t)nz_cap + 4 * 2 * TS + 15) & ~(size_t)15;
  return (int)(wave + 16 * 64 * MAXP + MAXSLOT * slot);
}

__global__ __launch_bounds__(64) void cell_search(LaunchDev P, const TaskDev* __restrict__ tasks) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  WaveLds& W = *reinterpret_cast<WaveLds*>(lds);
  const size_t wave_bytes = (sizeof(WaveLds) + 15) & ~(size_t)15;
  const size_t slot_bytes = (8 * 5 * (size_t)P.nz_cap + 4 * 2 * TS + 15) & ~(size_t)15;
  double2* const terms = reinterpret_cast<double2*>(lds + wave_bytes);
  const int lane = threadIdx.x;
  const int64_t wid = blockIdx.x;
  float* ring_base = P.ring + wid * (int64_t)MAXSLOT * P.ring_cap;
  float2* hist_base = P.hist + wid * (int64_t)MAXSLOT * HIST;
  const double TWO_PI = mcm::TWO_PI;
  const double ux = 1.0, uy = 0.0, vx = -0.0, vy = 1.0;  // chatterAt's feed direction

  while (true) {
    if (lane == 0) W.task = atomicAdd(P.queue, 1);
    wave_sync();
    const int task = uni(W.task);
    if (task >= P.ntasks) break;
    const TaskDev T = tasks[task];
    const int cell = uni(T.cell);
    const CellDev& C = P.cells[cell];
    const int N = C.N, spt = C.spt, spr = C.spr, decim = C.decim, nmx = C.nmx, nmy = C.nmy;
    const double fz = C.fz, dphi = C.dphi, pitch = C.pitch, binScale = C.binScale, hMax = C.hMax;
    const double dt = C.dt, Ktc = C.Ktc, Krc = C.Krc, Kte = C.Kte, Kre = C.Kre, mcx = C.mc;
    const long long steps = (long long)spt * N * C.revs;
    for (int j = lane; j < MC_BINS; j += 64) W.Hb[j] = C.Hb[j];

    // ---- bisection state (wave-uniform)
    const bool amax_task = T.mode == TASK_SINGLE;   // a single trial at T.ap (the aMax check)
    const bool check_ends = T.mode == TASK_BRACKET;
    const int iters = T.iters;
    const double tol = T.tol;
    double lo = T.lo, hi = T.hi;
    int it = 0;
    bool pend_lo = check_ends, pend_hi = check_ends, failed = false, last_chat = false;
    long long p_trials = 0, p_steps = 0, p_visits = 0, p_eng = 0;  // along the serial path
    long long x_trials = 0;                                        // executed
    double last_index = 0, last_amp = 0, last_once = 0;
    int last_nz = 0;

    while (true) {
      // ================= choose this round's trials and the slot size
      int nnode = 0, L = 0, SG = 64;
      bool ev_lo = false, ev_hi = false;
      if (amax_task) {
        W.node_ap[0] = T.ap;  // uniform write, identical in every lane
        nnode = 1;
      } else {
        if (!pend_lo && !pend_hi && !(it < iters && hi - lo > tol)) break;
        double dzz;
        const int nv_lo = pend_lo ? N * slices_at(C.toolFluteLen, lo, dzz) : 0;
        const int nv_hi = pend_hi ? N * slices_at(C.toolFluteLen, hi, dzz) : 0;
        // hint ends first; the tree is speculated only once no end stays pending
        for (int mode = 0; mode < 3 && nnode == 0; mode++) {
          const bool use_lo = pend_lo, use_hi = pend_hi && (mode < 2 || !pend_lo);
          if (mode == 2 && !(pend_lo && pend_hi)) break;
          const bool lo_only = mode == 2;  // the two ends one round each
          const int ne = lo_only ? 1 : (int)use_lo + (int)use_hi;
          const bool tree_ok = !lo_only && (use_lo == pend_lo) && (use_hi == pend_hi);
          for (int cand = (tree_ok ? 3 : 0); cand >= 0; cand--) {
            if (mode == 1 && cand > 0) continue;  // second try: ends alone
            const int tree = (1 << cand) - 1;
            const int need = ne + tree;
            if (need == 0) continue;
            int slots = 1;
            while (slots < need) slots <<= 1;
            if (slots > MAXSLOT) continue;
            const int sg = 64 / slots;
            // node intervals in breadth-first order: children of j are 2j+1 (chatter at
            // the midpoint: bracket [lo, m]) and 2j+2 (stable: [m, hi])
            double nl[MAXSLOT], nh[MAXSLOT], na[MAXSLOT];
            nl[0] = lo; nh[0] = hi;
            for (int j = 0; j < tree; j++) {
              const double m = __builtin_sqrt(nl[j] * nh[j]);
              na[j] = m;
              if (2 * j + 2 < tree) { nl[2 * j + 1] = nl[j]; nh[2 * j + 1] = m; nl[2 * j + 2] = m; nh[2 * j + 2] = nh[j]; }
            }
            int nvmax = lo_only ? nv_lo : max(use_lo ? nv_lo : 0, use_hi ? nv_hi : 0);
            for (int j = 0; j < tree; j++) nvmax = max(nvmax, N * slices_at(C.toolFluteLen, na[j], dzz));
            if (nvmax > MAXP * sg || nvmax > P.nv_cap) continue;
            L = cand; SG = sg; nnode = need;
            int q = 0;
            if (use_lo) { W.node_ap[q++] = lo; ev_lo = true; }
            if (use_hi && !lo_only) { W.node_ap[q++] = hi; ev_hi = true; }
            for (int j = 0; j < tree; j++) { W.node_ap[q + j] = na[j]; W.node_lo[q + j] = nl[j]; W.node_hi[q + j] = nh[j]; }
            break;
          }
        }
        if (nnode == 0) { failed = true; break; }  // the host sizes the trees so this never happens
      }
      wave_sync();
      const int nslot = nnode;
      const int slot = lane / SG, g = lane % SG;
      const bool in_slot = slot < nslot;
      SlotLds S;
      {
        const int sl = in_slot ? slot : 0;
        S.term = terms + sl * SG * MAXP;
        unsigned char* sb = lds + wave_bytes + 16 * 64 * MAXP + (size_t)sl * slot_bytes;
        double* d = reinterpret_cast<double*>(sb);
        S.sLag = d; d += P.nz_cap; S.sZ = d; d += P.nz_cap; S.sB = d; d += P.nz_cap;
        S.sKap = d; d += P.nz_cap; S.sR = d; d += P.nz_cap;
        S.tsx = reinterpret_cast<float*>(d); S.tsy = S.tsx + TS;
      }
      float* ring = ring_base + (int64_t)(in_slot ? slot : 0) * P.ring_cap;
      float2* hist = hist_base + (in_slot ? slot : 0) * HIST;

      // ---- fresh engine per slot at its depth (configure + setEngagement)
      const double ap = in_slot ? W.node_ap[slot] : 0.0;
      double dz = 0.25;
      const int nz = in_slot ? slices_at(C.toolFluteLen, ap, dz) : 0;
      const int NV = N * nz;
      const double db = dz;
      const double pdK = C.Ci * db * 1e-3 / (C.omega * 1e-3);
      for (int k = g; k < nz; k += SG) {
        const double b = js_min(dz, ap - k * dz);
        const double zr = k * dz + 0.5 * b;
        double r = C.R, kap = 1;
        if (C.ball && zr < C.R) {
          const double t = C.R - zr;
          r = __builtin_sqrt(js_max(1e-6, C.R * C.R - t * t));
          kap = r / C.R;
        }
        S.sB[k] = b; S.sZ[k] = -ap + zr; S.sR[k] = r; S.sKap[k] = kap;
        S.sLag[k] = zr * C.tanB / C.R;
      }
      for (int j = g; j < nz * spr; j += SG) ring[j] = 0.0f;
      if (g == 0 && in_slot) W.eng[slot] = 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();

      // per-pass visit of this lane: tooth i, slice k
      int vi[MAXP], vk[MAXP];
      bool vok[MAXP];
#pragma unroll
      for (int p = 0; p < MAXP; p++) {
        const int v = g + p * SG;
        vok[p] = in_slot && v < NV;
        vi[p] = vok[p] ? v / nz : 0;
        vk[p] = vok[p] ? v - vi[p] * nz : 0;
      }

      // modal state of each slot in LDS; the kick vx = 0.01 enters the first mode
      if (g == 0 && in_slot) {
#pragma unroll
        for (int j = 0; j < 2 * MC_MAX_MODES; j++) { W.mq[slot][j] = 0; W.mv[slot][j] = 0; }
        W.mv[slot][0] = 0.0 + (0.01 - 0.0);
      }
      double x = 0, y = 0, vxs = 0.01, vys = 0;
      int pr = 0, ph = 0, dc = 0, ti = 0, tcount = 0, hpos = 0, hcount = 0;
      long long my_eng = 0;
      float cur[MAXP];
#pragma unroll
      for (int p = 0; p < MAXP; p++) {
        cur[p] = 0.0f;
        if (vok[p]) {
          double phi = mcm::fmod_2pi((0.0 * dphi + (double)vi[p] * pitch) - S.sLag[vk[p]]);
          if (phi < 0) phi += TWO_PI;
          double sn, cs;
          mcm::sincos_js(phi, sn, cs);
          W.geo[p][lane] = make_double2(sn, cs);
          W.gbin[p][lane] = (int)(phi * binScale);
        }
      }
      wave_sync();

      for (long long s = 0; s < steps; s++) {
        const double xm = x * 1e3, ym = y * 1e3;  // rigid workpiece
        const double rvx = vxs, rvy = vys;
        const int prn = pr + 1 == spr ? 0 : pr + 1;
        const double thn = (double)prn * dphi;
        float nxt[MAXP];
#pragma unroll
        for (int p = 0; p < MAXP; p++) {
          nxt[p] = 0.0f;
          if (vok[p]) {
            int pi = prn + vi[p] * spt;
            if (pi >= spr) pi -= spr;
            nxt[p] = ring[pi * nz + vk[p]];
          }
        }
#pragma unroll
        for (int p = 0; p < MAXP; p++) {
          if (vok[p]) {
            const int i = vi[p], k = vk[p];
            const double2 gg = W.geo[p][lane];
            const double sn = gg.x, cs = gg.y;
            const double nxg = sn * ux + cs * vx, nyg = sn * uy + cs * vy;
            const double rn = xm * nxg + ym * nyg + C.dR[i];
            int pi = pr + i * spt;
            if (pi >= spr) pi -= spr;
            const double prev = (double)cur[p];
            const double feedH = fz * sn * S.sKap[k];
            const int bin = W.gbin[p][lane];
            const bool below = bin < MC_BINS ? ((double)W.Hb[bin] <= S.sZ[k]) : false;
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
                const double bk = S.sB[k];
                const double Ft = Ktc * se * h * bk + Kte * bk;
                double Fr = Krc * se * h * bk + Kre * bk;
                if (pdK != 0 && pdK == pdK) Fr += pdK * (bk / db) / S.sR[k] * (rvx * nxg + rvy * nyg);
                const double txg = cs * ux - sn * vx, tyg = cs * uy - sn * vy;
                tX = -Ft * txg - Fr * nxg;
                tY = -Ft * tyg - Fr * nyg;
              }
            }
            ring[pi * nz + k] = store;
            S.term[g + p * SG] = make_double2(tX, tY);
            // geometry of the next step (independent of the vibration)
            double phi = mcm::fmod_2pi((thn + (double)i * pitch) - S.sLag[k]);
            if (phi < 0) phi += TWO_PI;
            double nsn, ncs;
            mcm::sincos_js(phi, nsn, ncs);
            W.geo[p][lane] = make_double2(nsn, ncs);
            W.gbin[p][lane] = (int)(phi * binScale);
          }
        }
        wave_sync();
        double Fx = 0, Fy = 0;
        for (int v = 0; v < NV; v++) {
          const double2 t = S.term[v];
          Fx += t.x;
          Fy += t.y;
        }
        wave_sync();
        x = 0; y = 0; vxs = 0; vys = 0;
        if (in_slot) {
          double* mq = W.mq[slot];
          double* mv = W.mv[slot];
#pragma unroll
          for (int j = 0; j < MC_MAX_MODES; j++) {
            if (j < nmx) {
              double v = mv[j], q = mq[j];
              v += (Fx - C.cx[j] * v - C.kx[j] * q) / C.mx[j] * dt;
              q += v * dt;
              x += q;
              vxs += v;
              if (g == 0) { mv[j] = v; mq[j] = q; }
            }
          }
#pragma unroll
          for (int j = 0; j < MC_MAX_MODES; j++) {
            if (j < nmy) {
              double v = mv[MC_MAX_MODES + j], q = mq[MC_MAX_MODES + j];
              v += (Fy - C.cy[j] * v - C.ky[j] * q) / C.my[j] * dt;
              q += v * dt;
              y += q;
              vys += v;
              if (g == 0) { mv[MC_MAX_MODES + j] = v; mq[MC_MAX_MODES + j] = q; }
            }
          }
        }
        pr = prn;
        ph = ph + 1 == spt ? 0 : ph + 1;
        if (ph == 0) {
          if (g == 0 && in_slot) { S.tsx[ti] = (float)x; S.tsy[ti] = (float)y; }
          ti = (ti + 1) & (TS - 1);
          tcount++;
        }
        dc = dc + 1 == decim ? 0 : dc + 1;
        if (dc == 0) {
          if (g == 0 && in_slot) hist[hpos] = make_float2((float)(x * 1e6), (float)(y * 1e6));
          hpos = (hpos + 1) & (HIST - 1);
          hcount++;
        }
#pragma unroll
        for (int p = 0; p < MAXP; p++) cur[p] = nxt[p];
        wave_sync();
      }

      // ---- chatter index per slot (src/cutting.js:299-318)
      if (in_slot) atomicAdd(reinterpret_cast<unsigned long long*>(&W.eng[slot]), (unsigned long long)my_eng);
      if (g == 0 && in_slot) {
        double index = 0, amp = 0, once = 0;
        const int n = tcount < 24 ? tcount : 24;
        if (n >= 8) {
          double mx = 0, my = 0, v = 0;
          for (int k = 0; k < n; k++) { const int j = (ti - 1 - k) & (TS - 1); mx += S.tsx[j]; my += S.tsy[j]; }
          mx /= n; my /= n;
          for (int k = 0; k < n; k++) {
            const int j = (ti - 1 - k) & (TS - 1);
            const double dx = S.tsx[j] - mx, dy = S.tsy[j] - my;
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
        W.idx[slot] = index; W.amp[slot] = amp; W.once[slot] = once; W.nz[slot] = nz;
      }
      wave_sync();
      x_trials += nslot;

      // ---- walk the tree with the verdicts (wave-uniform)
      auto take = [&](int j) {
        p_trials++;
        p_steps += steps;
        p_visits += (long long)N * W.nz[j] * steps;
        p_eng += W.eng[j];
        last_index = W.idx[j]; last_amp = W.amp[j]; last_once = W.once[j]; last_nz = W.nz[j];
        last_chat = W.idx[j] > MC_CHATTER_IDX;
      };
      if (amax_task) {
        take(0);
        break;
      }
      int q = 0;
      if (ev_lo) {
        pend_lo = false;
        take(q++);                     // chatterAt(lo): must be stable
        if (last_chat) { failed = true; break; }
      }
      if (ev_hi) {
        pend_hi = false;
        take(q++);                     // chatterAt(hi): must chatter
        if (!last_chat) { failed = true; break; }
      }
      if (pend_lo || pend_hi) { wave_sync(); continue; }
      int j = 0;
      for (int l = 0; l < L; l++) {
        if (!(it < iters && hi - lo > tol)) break;
        take(q + j);
        const double m = W.node_ap[q + j];
        if (last_chat) { hi = m; j = 2 * j + 1; } else { lo = m; j = 2 * j + 2; }
        it++;
      }
      wave_sync();
    }

    if (lane == 0) {
      ResultDev& R = P.out[task];
      R.alim = __builtin_sqrt(lo * hi);
      R.lo = lo; R.hi = hi;
      R.index = last_index; R.amp = last_amp; R.once = last_once;
      R.open = 0; R.trials = (int)p_trials; R.chatter = last_chat; R.nz = last_nz;
      R.fell_back = failed; R.pad = (int)x_trials;
      R.steps = p_steps; R.visits = p_visits; R.engaged = p_eng;
    }
    wave_sync();
  }
}

int mcd_launch_search(const LaunchDev& P, const TaskDev* d_tasks, int blocks, hipStream_t stream) {
  const size_t lds = (size_t)mcd_search_lds_bytes(P.nv_cap, P.nz_cap);
  hipLaunchKernelGGL(cell_search, dim3(blocks), dim3(64), lds, stream, P, d_tasks);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
