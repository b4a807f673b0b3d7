// search.hip — speculative, wave-per-cell time-domain stability search for gfx950.
//
// timeDomainLimit (src/stabtd.js:24-32) checks the depth aMax, then bisects
// geometrically: every trial is a full chatterAt simulation (src/cutting.js:146-318) of
// spt*N*revs steps whose next depth depends on the previous verdict.  Here one wavefront
// owns one cell and, per ROUND, evaluates the next L levels of the bisection tree at once
// (1, 3 or 7 trials, one per lane SLOT of 64, 16 or 8 lanes); it then walks the tree with
// the verdicts, which reproduces the serial bisection exactly (same depths, same
// brackets) in ceil(levels / L) rounds.  All trials of a round have the same length (same
// cell), so the slots advance in lockstep, and all cell constants are wave-uniform.  The
// aMax check is a separate task of its own wave, run concurrently; the host combines the
// two (stable at aMax -> open result).
//
// Per step the lanes of a slot split the N*nz (tooth, slice) visits (at most MAXP passes;
// the slot size is chosen per round from the trial depths), leave their force terms in
// LDS, then every lane sums its slot's terms serially in the reference's (tooth, slice)
// order (bit-identical FP64).  The modes of a slot are integrated in parallel, one lane
// per mode (one division for all modes), and summed in mode order.
//
// Edge geometry table: the edge angle of tooth i on slice k at step pr,
// phi = (pr*dphi + i*pitch - lag_k) mod 2pi, depends only on (pr, i, lag_k) and repeats
// every revolution, so its sine, cosine and engagement-map bin are tabulated once per cell
// (full slices) and once per round (the partial top slice of each trial) in a per-wave
// global table; the step loop reads it (prefetched a step ahead, like the delay line)
// instead of evaluating fdlibm sin/cos per visit.  A round needing more distinct lags than
// the table holds (possible only when the depth changes the slice pitch) aborts the task
// with fell_back = 2, and the host runs that cell on the one-trial-per-group kernel.
//
// Wide tasks: a workgroup is four waves.  It first takes the wide tasks (the trees of the
// cells with the longest trials) one at a time with all four waves, up to 2 slots per wave:
// the round plan and the verdicts live in wave 0's task block, every wave simulates its
// own slots, and the waves meet at a barrier to walk the tree, so deep trials, which leave
// room for one or two slots in a single wave, get 8 slots of 32 lanes.  Once no wide task
// is left, each wave takes ordinary tasks on its own (up to 8 slots per wave).
#include <hip/hip_runtime.h>

#include "mc_dev.h"
#include "mc_math.h"

using namespace mcd;

namespace {

constexpr int MAXP = 2;        // passes of a slot's lanes per step
constexpr int MAXNODE = 8;     // trials of a task per round (over all its waves)
constexpr int MAXSPEC = 16;    // per-round table columns for partial / non-standard slices

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ void global_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double js_min(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double js_max(double a, double b) { return a > b ? a : b; }

// slice count of chatterAt's engine at depth ap (src/stabtd.js:16, cutting.js:63-64,101)
__device__ __forceinline__ int slices_at(double toolFL, double ap, double& dz) {
  const double fl = js_max(toolFL, ap + 1);
  dz = js_max(0.25, fl / 48);
  const int maxSlices = (int)__builtin_ceil(fl / dz);
  return ap <= 0 ? 0 : (int)js_min((double)maxSlices, __builtin_ceil(ap / dz));
}

// the cell constants the step loop reads, staged in LDS
struct CellCst {
  double fz, dphi, pitch, binScale, hMax, dt, Ktc, Krc, Kte, Kre, mc;
  double dR[MC_MAX_FLUTES];
};

// per wave: cell constants, the modal state, samples and engagement map of its slots
template <int SPW>
struct WaveLds {
  CellCst K;
  double mq[SPW][2 * MC_MAX_MODES], mv[SPW][2 * MC_MAX_MODES];
  double spec_lag[MAXSPEC];
  float Hb[MC_BINS];
  float tsx[SPW][TS], tsy[SPW][TS];  // tooth-passing samples of each slot
  int nspec, overflow;
};

// per task (workgroup): the task index, the round plan (bisection-tree nodes in best-first
// order, their intervals and the indices of their evaluated children, -1 if none), the
// best-first frontier (thread 0's working storage) and the verdicts of the round's trials
struct TaskLds {
  int task;
  double node_ap[MAXNODE], node_lo[MAXNODE], node_hi[MAXNODE];
  int node_kid[MAXNODE][2];
  int plan_n, plan_sg, plan_spw, plan_first, plan_ntree, plan_evlo, plan_evhi;
  double fr_lo[2 * MAXNODE + 2], fr_hi[2 * MAXNODE + 2], fr_p[2 * MAXNODE + 2];
  int fr_par[2 * MAXNODE + 2], fr_side[2 * MAXNODE + 2], fr_d[2 * MAXNODE + 2];
  double idx[MAXNODE], amp[MAXNODE], once[MAXNODE];
  long long eng[MAXNODE];
  int nz[MAXNODE];
  int overflow[8];
};

// Behind WaveLds the wave's LDS holds, at different times of a round, either the slice
// records of every slot (set-up: lag, z, width, kappa, radius, table column) or the step
// loop's buffers: the engaged visits of each slot (h, sin, cos, width, damping factor),
// then their force terms in (tooth, slice) order.
struct SlotLds {
  double* sLag;    // [nz_cap] per slice
  double* sZ;
  double* sB;
  double* sKap;
  double* sR;
  int* lagid;      // table column of each slice
};
constexpr int NREC = 64 * MAXP;  // engaged-visit records of a wave (a slot of SG lanes owns SG*MAXP)

__device__ __host__ __forceinline__ size_t task_bytes() { return (sizeof(TaskLds) + 15) & ~(size_t)15; }
template <int SPW>
__device__ __host__ __forceinline__ size_t wave_bytes() { return (sizeof(WaveLds<SPW>) + 15) & ~(size_t)15; }
// a slot's set-up records: up to MAXP*SG slices (a trial's N*nz visits fit its MAXP*SG
// lane passes), 5 doubles and 1 int each; the slots of a wave together stay within the
// step buffers (MAXP*64 slices)
__device__ __forceinline__ size_t slot_stride(int SG) { return (size_t)MAXP * SG * 44; }
__device__ __host__ __forceinline__ size_t step_bytes() { return (size_t)NREC * (16 + 5 * 8); }
// a wave's LDS: its task block, its wave block, then its step buffers (which the set-up
// records of its slots share: 64*MAXP*44 bytes < step_bytes())
__device__ __host__ __forceinline__ size_t per_wave_bytes() {
  return task_bytes() + wave_bytes<8>() + step_bytes();
}
static_assert(64 * MAXP * 44 <= NREC * (16 + 5 * 8), "set-up records fit the step buffers");

template <int WW>
__device__ __forceinline__ void task_sync() {
  if (WW > 1) __syncthreads();
  else wave_sync();
}

// one table entry: sin, cos and bin of the edge angle at (pr, tooth i, lag)
__device__ __forceinline__ void edge_geometry(double theta, double ipitch, double lag, double binScale,
                                              double& sn, double& cs, int& bin) {
  double phi = mcm::fmod_2pi((theta + ipitch) - lag);
  if (phi < 0) phi += mcm::TWO_PI;
  mcm::sincos_js(phi, sn, cs);
  bin = (int)(phi * binScale);
}

// P(ln a < ln alim < ln b) under N(mu, sig^2), with tails taken from erfc for precision
__device__ __forceinline__ double prior_mass(double a, double b, double mu, double sig) {
  const double xa = (::log(a) - mu) / (sig * 1.4142135623730951);
  const double xb = (::log(b) - mu) / (sig * 1.4142135623730951);
  if (xa > 0) return 0.5 * (erfc(xa) - erfc(xb));
  if (xb < 0) return 0.5 * (erfc(-xb) - erfc(-xa));
  return 1.0 - 0.5 * erfc(xb) - 0.5 * erfc(-xa);
}

// Up to kmax nodes of the bisection tree below [lo, hi], best first by the probability
// that the serial bisection visits them (the prior mass of the node's interval; without a
// prior, 2^-depth, i.e. breadth first).  Only nodes the serial search could still evaluate
// (iteration budget and stopping width) are taken.  Any prefix of the order is a subtree
// containing the root.  Runs in one lane; returns the node count.
__device__ int plan_tree(TaskLds& W, double lo, double hi, int it, int iters, double tol, double mu, double sig,
                         int kmax) {
  int nf = 0, n = 0;
  W.fr_lo[0] = lo; W.fr_hi[0] = hi; W.fr_p[0] = 1.0; W.fr_par[0] = -1; W.fr_side[0] = 0; W.fr_d[0] = 0;
  nf = 1;
  while (n < kmax && nf > 0) {
    int b = 0;
    for (int f = 1; f < nf; f++)
      if (W.fr_p[f] > W.fr_p[b] || (W.fr_p[f] == W.fr_p[b] && W.fr_d[f] < W.fr_d[b])) b = f;
    const double a = W.fr_lo[b], c = W.fr_hi[b], m = __builtin_sqrt(a * c);
    const int d = W.fr_d[b], par = W.fr_par[b], side = W.fr_side[b];
    W.node_lo[n] = a; W.node_hi[n] = c; W.node_ap[n] = m;
    W.node_kid[n][0] = W.node_kid[n][1] = -1;
    if (par >= 0) W.node_kid[par][side] = n;
    nf--;
    W.fr_lo[b] = W.fr_lo[nf]; W.fr_hi[b] = W.fr_hi[nf]; W.fr_p[b] = W.fr_p[nf];
    W.fr_par[b] = W.fr_par[nf]; W.fr_side[b] = W.fr_side[nf]; W.fr_d[b] = W.fr_d[nf];
    // children: chatter at m -> [a, m], stable -> [m, c]
    for (int sd = 0; sd < 2; sd++) {
      const double ca = sd ? m : a, cc = sd ? c : m;
      if (!(it + d + 1 < iters && cc - ca > tol)) continue;
      W.fr_lo[nf] = ca; W.fr_hi[nf] = cc;
      W.fr_p[nf] = sig > 0 ? prior_mass(ca, cc, mu, sig) : __builtin_ldexp(1.0, -(d + 1));
      W.fr_par[nf] = n; W.fr_side[nf] = sd; W.fr_d[nf] = d + 1;
      nf++;
    }
    n++;
  }
  return n;
}

}  // namespace

// LDS of a workgroup: per wave a task block, a wave block and the step buffers
int mcd_search_lds_bytes(int nz_cap) {
  (void)nz_cap;
  return (int)(4 * per_wave_bytes());
}
int mcd_search_maxspec() { return MAXSPEC; }

// one task: WW waves (wv = this wave's index among them), up to SPW slots per wave
template <int WW, int SPW>
__device__ __forceinline__ void run_task(const LaunchDev& P, const TaskDev* __restrict__ tasks,
                                         ResultDev* __restrict__ outp, int task, TaskLds& TL,
                                         WaveLds<8>& W, unsigned char* const area, const int wv,
                                         const int lane, float* const ring_base, float2* const hist_base,
                                         double2* const gsc, unsigned char* const gbn) {
  double2* const terms = reinterpret_cast<double2*>(area);
  double* const rec_h = reinterpret_cast<double*>(area + 16 * NREC);
  double* const rec_sn = rec_h + NREC;
  double* const rec_cs = rec_sn + NREC;
  double* const rec_b = rec_cs + NREC;
  double* const rec_pd = rec_b + NREC;
  const char* const gsc_b = reinterpret_cast<const char*>(gsc);
  char* const ring_b = reinterpret_cast<char*>(ring_base);
  const int NL = P.nz_cap + MAXSPEC;   // table columns: full slices, then per-round ones
  const bool lead = lane == 0 && wv == 0;  // plans the rounds and writes the result
  {
    const long long t_start = wall_clock64();
    const TaskDev T = tasks[task];
    const int cell = uni(T.cell);
    const CellDev& C = P.cells[cell];
    const int N = C.N, spt = C.spt, spr = C.spr, decim = C.decim;
    const int steps = spt * N * C.revs;  // below 2^31: the host bounds spt, N and revs
    for (int j = lane; j < MC_BINS; j += 64) W.Hb[j] = C.Hb[j];
    if (lane == 0) {
      CellCst& K = W.K;
      K.fz = C.fz; K.dphi = C.dphi; K.pitch = C.pitch; K.binScale = C.binScale; K.hMax = C.hMax;
      K.dt = C.dt; K.Ktc = C.Ktc; K.Krc = C.Krc; K.Kte = C.Kte; K.Kre = C.Kre; K.mc = C.mc;
      for (int j = 0; j < MC_MAX_FLUTES; j++) K.dR[j] = C.dR[j];
    }
    wave_sync();
    const CellCst& K = W.K;
    // standard slice pitch of this cell (depths with ap + 1 <= flute length)
    const double dz0 = js_max(0.25, C.toolFluteLen / 48);
    int kf_built = 0;  // full-slice table columns built for this task

    // ---- bisection state (wave-uniform)
    const bool amax_task = T.mode == TASK_SINGLE;  // a single trial at T.ap (the aMax check)
    const bool check_ends = T.mode == TASK_BRACKET;
    const int iters = T.iters;
    const double tol = T.tol;
    double lo = T.lo, hi = T.hi;
    int it = 0;
    bool pend_lo = check_ends, pend_hi = check_ends, failed = false, last_chat = false;
    bool tab_overflow = false;  // more distinct slice lags than table columns: host reruns the cell
    long long p_trials = 0, p_steps = 0, p_visits = 0, p_eng = 0;  // along the serial path
    long long x_trials = 0;                                        // executed
    double last_index = 0, last_amp = 0, last_once = 0;
    int last_nz = 0;

    while (true) {
      // ================= this round's trials and slot size (planned by thread 0)
      if (!amax_task && !pend_lo && !pend_hi && !(it < iters && hi - lo > tol)) break;
      if (lead) {
        // slot layouts: S slots over the task's waves, max(1, S / WW) per wave
        int nn = 0, sgp = 64, spwp = 1, first = 0, ntree = 0, elo = 0, ehi = 0;
        double dzz;
        if (amax_task) {
          TL.node_ap[0] = T.ap;
          nn = 1;
        } else if (pend_lo || pend_hi) {
          // hint ends first (both, else the high one alone, else the low one alone), with
          // the first tree levels below the hint speculated where lanes allow
          const int nv_lo = pend_lo ? N * slices_at(C.toolFluteLen, lo, dzz) : 0;
          const int nv_hi = pend_hi ? N * slices_at(C.toolFluteLen, hi, dzz) : 0;
          for (int mode = 0; mode < 3 && nn == 0; mode++) {
            const bool use_lo = pend_lo && mode != 1, use_hi = pend_hi && mode != 2;
            if (mode > 0 && !(pend_lo && pend_hi)) break;
            const int ne = (int)use_lo + (int)use_hi;
            for (int ki = (mode == 0 ? 0 : 2); ki < 3; ki++) {
              const int k = ki == 0 ? 6 : ki == 1 ? 2 : 0;   // tree nodes beside the ends
              const int need = ne + k;
              int slots = 1;
              while (slots < need) slots <<= 1;
              if (slots > WW * SPW) continue;
              const int spw = slots > WW ? slots / WW : 1, sg = 64 / spw;
              const int nt = k ? plan_tree(TL, lo, hi, it, iters, tol, T.mu, T.sig, k) : 0;
              int nvmax = max(use_lo ? nv_lo : 0, use_hi ? nv_hi : 0);
              for (int j = 0; j < nt; j++) nvmax = max(nvmax, N * slices_at(C.toolFluteLen, TL.node_ap[j], dzz));
              if (nvmax <= MAXP * sg && nvmax <= P.nv_cap) {
                // tree nodes move behind the ends
                for (int j = nt - 1; j >= 0; j--) {
                  TL.node_ap[ne + j] = TL.node_ap[j]; TL.node_lo[ne + j] = TL.node_lo[j]; TL.node_hi[ne + j] = TL.node_hi[j];
                  TL.node_kid[ne + j][0] = TL.node_kid[j][0]; TL.node_kid[ne + j][1] = TL.node_kid[j][1];
                }
                int q = 0;
                if (use_lo) { TL.node_ap[q++] = lo; elo = 1; }
                if (use_hi) { TL.node_ap[q++] = hi; ehi = 1; }
                nn = need; sgp = sg; spwp = spw; first = ne; ntree = nt;
                break;
              }
            }
          }
        } else {
          // the largest best-first prefix of the tree whose trials fit the lanes
          const int nt = plan_tree(TL, lo, hi, it, iters, tol, T.mu, T.sig, WW * SPW);
          for (int slots = WW * SPW; slots >= 1 && nn == 0; slots >>= 1) {
            const int k = min(slots, nt);
            const int spw = slots > WW ? slots / WW : 1, sg = 64 / spw;
            int nvmax = 0;
            for (int j = 0; j < k; j++) nvmax = max(nvmax, N * slices_at(C.toolFluteLen, TL.node_ap[j], dzz));
            if (k >= 1 && nvmax <= MAXP * sg && nvmax <= P.nv_cap) { nn = k; sgp = sg; spwp = spw; ntree = k; }
          }
        }
        // a tree whose cell's aMax check came out stable is moot: the cell is open and the
        // host reads only the check (src/stabtd.js:25), so the tree stops here
        if (T.mode == TASK_BISECT && P.cstate &&
            __hip_atomic_load(P.cstate + cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 2)
          nn = -1;
        TL.plan_n = nn; TL.plan_sg = sgp; TL.plan_spw = spwp; TL.plan_first = first; TL.plan_ntree = ntree;
        TL.plan_evlo = elo; TL.plan_evhi = ehi;
      }
      task_sync<WW>();
      const int nnode = uni(TL.plan_n), SG = uni(TL.plan_sg), spw = uni(TL.plan_spw);
      const int tfirst = uni(TL.plan_first), ntree = uni(TL.plan_ntree);
      const bool ev_lo = uni(TL.plan_evlo) != 0, ev_hi = uni(TL.plan_evhi) != 0;
      if (nnode < 0) break;                       // moot (open cell)
      if (nnode == 0) { failed = true; break; }  // the host sizes the trees so this never happens
      // this wave runs the task's slots w*spw .. w*spw + spw - 1 (lslot is the local index)
      const int lslot = lane / SG, g = lane % SG;
      const int gslot = wv * spw + lslot;
      const bool in_slot = lslot < spw && gslot < nnode;
      const int sl = in_slot ? lslot : 0;
      const int nloc = max(0, min(spw, nnode - wv * spw));   // this wave's slots
      SlotLds S;
      {
        double* d = reinterpret_cast<double*>(area + (size_t)sl * slot_stride(SG));
        const int cap = MAXP * SG;
        S.sLag = d; d += cap; S.sZ = d; d += cap; S.sB = d; d += cap;
        S.sKap = d; d += cap; S.sR = d; d += cap;
        S.lagid = reinterpret_cast<int*>(d);
      }
      const int rbase = sl * SG * MAXP;  // this slot's visit records and force terms
      double2* const term = terms + rbase;
      float* const tsx = W.tsx[sl];
      float* const tsy = W.tsy[sl];
      float* ring = ring_base + (int64_t)sl * P.ring_cap;
      float2* hist = hist_base + sl * HIST;

      // ---- fresh engine per slot at its depth (configure + setEngagement)
      const double ap = in_slot ? TL.node_ap[gslot] : 0.0;
      double dz = dz0;
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
      wave_sync();

      // ---- table columns: full slices k carry lag(k*dz0 + dz0/2); the other slices get
      // per-round columns (deduplicated bit for bit) or, past MAXSPEC, direct evaluation
      int kneed = 0;
      for (int s2 = 0; s2 < nloc; s2++) {
        double dd = dz0;
        kneed = max(kneed, slices_at(C.toolFluteLen, TL.node_ap[wv * spw + s2], dd));
      }
      kneed = min(kneed, P.nz_cap);
      if (lane == 0) {
        int ns = 0;
        for (int s2 = 0; s2 < nloc; s2++) {
          double* d = reinterpret_cast<double*>(area + (size_t)s2 * slot_stride(SG));
          const double* lag2 = d;
          int* id2 = reinterpret_cast<int*>(d + 5 * MAXP * SG);
          double dd = dz0;
          const int nz2 = slices_at(C.toolFluteLen, TL.node_ap[wv * spw + s2], dd);
          for (int k = 0; k < nz2; k++) {
            const double full = (k * dz0 + 0.5 * dz0) * C.tanB / C.R;
            int id = -1;
            if (__builtin_bit_cast(uint64_t, lag2[k]) == __builtin_bit_cast(uint64_t, full)) {
              id = k;
            } else {
              for (int q = 0; q < ns; q++)
                if (__builtin_bit_cast(uint64_t, W.spec_lag[q]) == __builtin_bit_cast(uint64_t, lag2[k])) id = P.nz_cap + q;
              if (id < 0 && ns < MAXSPEC) { W.spec_lag[ns] = lag2[k]; id = P.nz_cap + ns; ns++; }
            }
            id2[k] = id;
          }
        }
        W.nspec = ns;
        int ovf = 0;
        for (int s2 = 0; s2 < nloc; s2++) {
          double* d = reinterpret_cast<double*>(area + (size_t)s2 * slot_stride(SG));
          const int* id2 = reinterpret_cast<int*>(d + 5 * MAXP * SG);
          double dd = dz0;
          const int nz2 = slices_at(C.toolFluteLen, TL.node_ap[wv * spw + s2], dd);
          for (int k = 0; k < nz2; k++) if (id2[k] < 0) ovf = 1;
        }
        TL.overflow[wv] = ovf;
      }
      task_sync<WW>();
      const int nspec = uni(W.nspec);
      int any_ovf = 0;
      for (int v = 0; v < WW; v++) any_ovf |= TL.overflow[v];
      if (uni(any_ovf)) { failed = true; tab_overflow = true; break; }
      {
        // extend the full-slice columns, then (re)build this round's columns
        const int kf_new = max(kf_built, kneed);
        const long long pe = (long long)spr * N;
        for (int col = 0; col < 2; col++) {
          const int c0 = col == 0 ? kf_built : P.nz_cap, c1 = col == 0 ? kf_new : P.nz_cap + nspec;
          const int nc = c1 - c0;
          if (nc <= 0) continue;
          for (long long e = lane; e < pe * nc; e += 64) {
            const int c = c0 + (int)(e % nc);
            const long long pi_ = e / nc;
            const int pr = (int)(pi_ / N), i = (int)(pi_ % N);
            const double lag = col == 0 ? (c * dz0 + 0.5 * dz0) * C.tanB / C.R : W.spec_lag[c - P.nz_cap];
            double sn, cs;
            int bin;
            edge_geometry((double)pr * K.dphi, (double)i * K.pitch, lag, K.binScale, sn, cs, bin);
            const long long at = pi_ * NL + c;
            gsc[at] = make_double2(sn, cs);
            gbn[at] = (unsigned char)(bin < 255 ? bin : 255);
          }
        }
        kf_built = kf_new;
      }
      global_sync();

      // per-pass visit of this lane: tooth i, slice k, table column.  Lanes without a
      // visit keep in-bounds indices (tooth 0, slice 0, column 0) so that every lane issues
      // the same loads each step: the delay-line and table prefetch then stays in flight
      // across the step instead of being waited for at once.
      int vi[MAXP], vk[MAXP], vc[MAXP];
      bool vok[MAXP];
      double vKap[MAXP], vZ[MAXP], vB[MAXP], vPd[MAXP], vdR[MAXP];
#pragma unroll
      for (int p = 0; p < MAXP; p++) {
        const int v = g + p * SG;
        vok[p] = in_slot && v < NV;
        vi[p] = vok[p] ? v / nz : 0;
        vk[p] = vok[p] ? v - vi[p] * nz : 0;
        vc[p] = vok[p] ? S.lagid[vk[p]] : 0;
        if (vc[p] < 0) vc[p] = 0;
        vKap[p] = vok[p] ? S.sKap[vk[p]] : 1.0;
        vZ[p] = vok[p] ? S.sZ[vk[p]] : 0.0;
        vB[p] = vok[p] ? S.sB[vk[p]] : 0.0;
        // process-damping factor of the slice: pdK * (b / db) / r
        vPd[p] = vok[p] ? pdK * (vB[p] / db) / S.sR[vk[p]] : 0.0;
        vdR[p] = K.dR[vi[p]];
      }
      const double fz = K.fz, hMax = K.hMax, Ktc = K.Ktc, Krc = K.Krc, Kte = K.Kte, Kre = K.Kre;
      const double mcx = K.mc, dt = K.dt;
      const bool use_se = mcx != 0 && mcx == mcx;
      // x = 0.05 / max(h, 0.002) lies in [0.05 / hMax, 25]; the host sends this kernel only
      // cells with hMax <= 1e27 and pow_cell_ok(mc), where pow_cell equals pow_js
      const double my1 = mcm::clear_lo(mcx), myr = mcx - my1;
      const bool use_pd = pdK != 0 && pdK == pdK;
      // lanes of this lane's slot in a wave-wide ballot
      const unsigned long long smask = SG == 64 ? ~0ull : (((1ull << SG) - 1) << (sl * SG));

      // modes: lane g < MC_MAX_MODES integrates x-mode g, lane MC_MAX_MODES + j y-mode j;
      // the kick vx = 0.01 enters the first x mode (src/stabtd.js:19, cutting.js:164)
      const bool xmode = in_slot && g < C.nmx;
      const bool ymode = in_slot && g >= MC_MAX_MODES && g - MC_MAX_MODES < C.nmy;
      const int mj = g < MC_MAX_MODES ? g : g - MC_MAX_MODES;
      double mk = 1, mm = 1, mcd_ = 0, mq = 0, mv = 0;
      if (xmode) { mk = C.kx[mj]; mm = C.mx[mj]; mcd_ = C.cx[mj]; }
      if (ymode) { mk = C.ky[mj]; mm = C.my[mj]; mcd_ = C.cy[mj]; }
      if (xmode && mj == 0) mv = 0.0 + (0.01 - 0.0);
      const int nmx = C.nmx, nmy = C.nmy;
      double x = 0, y = 0, vxs = 0.01, vys = 0;
      int pr = 0, ph = 0, dc = 0, ti = 0, tcount = 0, hpos = 0, hcount = 0;
      // a history sample is stored one step late, beside the next step's delay-line stores:
      // a store issued at the end of a step held up the wait for the next step's prefetch
      bool hpend = false;
      int hidx = 0;
      float2 hval = make_float2(0.0f, 0.0f);
      long long my_eng = 0;
      float cur[MAXP];
      double2 csc[MAXP];
      int cbin[MAXP];
      const unsigned rowN = (unsigned)N * (unsigned)NL;  // table entries per step row
      // per pass: delay-line row base of the slot and slice, the ring position of the
      // visit's tooth at the current step ((pr + i*spt) mod spr), and its table column
      unsigned rs[MAXP], rp[MAXP], tcol[MAXP];
#pragma unroll
      for (int p = 0; p < MAXP; p++) {
        cur[p] = 0.0f;
        rs[p] = (unsigned)sl * (unsigned)P.ring_cap + (unsigned)vk[p];
        rp[p] = (unsigned)(vi[p] * spt);  // pr = 0, and i*spt < spr
        tcol[p] = (unsigned)vi[p] * (unsigned)NL + (unsigned)vc[p];
        csc[p] = *reinterpret_cast<const double2*>(gsc_b + 16u * tcol[p]);
        cbin[p] = gbn[tcol[p]];
      }

      wave_sync();  // the slice records are read; the step loop reuses their LDS
      const int wsteps = nloc > 0 ? steps : 0;  // a wave without slots this round skips the steps
      for (int s = 0; s < wsteps; s++) {
        const double xm = x * 1e3, ym = y * 1e3;  // rigid workpiece
        const double rvx = vxs, rvy = vys;
        const int prn = pr + 1 == spr ? 0 : pr + 1;
        // prefetch the next step's delay-line values and edge geometry (all lanes; 32-bit
        // offsets from the wave's base pointers)
        float nxt[MAXP];
        double2 nsc[MAXP];
        int nbin[MAXP];
        unsigned rpn[MAXP];
        const unsigned rowat = (unsigned)prn * rowN;
#pragma unroll
        for (int p = 0; p < MAXP; p++) {
          rpn[p] = rp[p] + 1 == (unsigned)spr ? 0u : rp[p] + 1;
          nxt[p] = *reinterpret_cast<const float*>(ring_b + 4u * (rs[p] + rpn[p] * (unsigned)nz));
          const unsigned at = rowat + tcol[p];
          nsc[p] = *reinterpret_cast<const double2*>(gsc_b + 16u * at);
          nbin[p] = gbn[at];
        }
        // engagement-map heights at both passes' bins, read together
        float hb[MAXP];
#pragma unroll
        for (int p = 0; p < MAXP; p++) hb[p] = W.Hb[cbin[p] < MC_BINS ? cbin[p] : 0];
        // 1) chip thickness and delay-line update of every visit; the engaged ones are
        //    recorded in (tooth, slice) order, the order in which the reference sums them
        int ne = 0;
        float st_v[MAXP];
#pragma unroll
        for (int p = 0; p < MAXP; p++) {
          // src/cutting.js:176-195, with its branches as selects.  chatterAt's feed frame is
          // u = (1, 0), v = (-0, 1) (src/stabtd.js:18), so the edge normal
          // (sn*ux + cs*vx, sn*uy + cs*vy) is (sn, cs) bit for bit: sn is never -0 (the edge
          // angle is never -0) and cs is never 0 (no double is a zero of cosine)
          const double sn = csc[p].x, cs = csc[p].y;
          const double rn = xm * sn + ym * cs + vdR[p];
          const double prev = (double)cur[p];
          const double feedH = fz * sn * vKap[p];
          const bool below = cbin[p] < MC_BINS && (double)hb[p] <= vZ[p];
          const bool cutting = !(below || sn <= 0);
          const double h = feedH + rn - prev;
          const bool thin = h <= 0;  // the edge passes above the surface it left
          const float store = (cutting && thin) ? (float)(prev - feedH) : (float)rn;
          const bool eng = vok[p] && cutting && !thin;
          st_v[p] = store;
          const unsigned long long m = __builtin_amdgcn_ballot_w64(eng) & smask;
          const int before = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
          if (eng) {
            const int r = rbase + ne + before;
            rec_h[r] = h; rec_sn[r] = sn; rec_cs[r] = cs; rec_b[r] = vB[p]; rec_pd[r] = vPd[p];
          }
          ne += __builtin_popcountll(m);
        }
        // the delay-line stores after every pass's chip: issued together, behind the next
        // step's prefetch, they keep the prefetch's loads from being waited for mid-step
#pragma unroll
        for (int p = 0; p < MAXP; p++)
          if (vok[p]) *reinterpret_cast<float*>(ring_b + 4u * (rs[p] + rp[p] * (unsigned)nz)) = st_v[p];
        if (hpend && g == 0 && in_slot) hist[hidx] = hval;
        hpend = false;
        my_eng += (g == 0) ? ne : 0;
        wave_sync();
        // 2) cutting forces of the engaged visits, one visit per lane (src/cutting.js:196-206)
        if (in_slot) {
          for (int r = g; r < ne; r += SG) {
            double h = rec_h[rbase + r];
            const double sn = rec_sn[rbase + r], cs = rec_cs[rbase + r];
            const double bk = rec_b[rbase + r];
            if (h > hMax) h = hMax;
            const double se = use_se ? mcm::pow_cell(0.05 / (h > 0.002 ? h : 0.002), mcx, my1, myr) : 1.0;
            const double Ft = Ktc * se * h * bk + Kte * bk;
            double Fr = Krc * se * h * bk + Kre * bk;
            // edge normal (sn, cs) and tangent (cs, -sn) in chatterAt's feed frame; engaged
            // visits have sn > 0, where cs*uy - sn*vy is -sn bit for bit
            if (use_pd) Fr += rec_pd[rbase + r] * (rvx * sn + rvy * cs);
            term[r] = make_double2(-Ft * cs - Fr * sn, Ft * sn - Fr * cs);
          }
        }
        // pad to a multiple of 4 with +0 terms: the running sum starts at +0 and can never
        // be -0 under round-to-nearest, so adding +0 leaves it unchanged
        const int ne4 = (ne + 3) & ~3;
        if (in_slot && g < ne4 - ne) term[ne + g] = make_double2(0.0, 0.0);
        wave_sync();
        // 3) the serial sum, in order
        double Fx = 0, Fy = 0;
        for (int v = 0; v < ne4; v += 4) {
          const double2 t0 = term[v], t1 = term[v + 1], t2 = term[v + 2], t3 = term[v + 3];
          Fx += t0.x; Fy += t0.y;
          Fx += t1.x; Fy += t1.y;
          Fx += t2.x; Fy += t2.y;
          Fx += t3.x; Fy += t3.y;
        }
        // 4) modal integration (semi-implicit Euler, src/cutting.js:247-255): each mode lane
        //    advances its mode, then every lane sums the slot's modes in order
        if (xmode || ymode) {
          const double F = xmode ? Fx : Fy;
          mv += (F - mcd_ * mv - mk * mq) / mm * dt;
          mq += mv * dt;
          W.mq[sl][g] = mq;
          W.mv[sl][g] = mv;
        }
        wave_sync();
        x = 0; y = 0; vxs = 0; vys = 0;
        if (in_slot) {
          double aq[2 * MC_MAX_MODES], av[2 * MC_MAX_MODES];
#pragma unroll
          for (int j = 0; j < 2 * MC_MAX_MODES; j++) { aq[j] = W.mq[sl][j]; av[j] = W.mv[sl][j]; }
#pragma unroll
          for (int j = 0; j < MC_MAX_MODES; j++)
            if (j < nmx) { x += aq[j]; vxs += av[j]; }
#pragma unroll
          for (int j = 0; j < MC_MAX_MODES; j++)
            if (j < nmy) { y += aq[MC_MAX_MODES + j]; vys += av[MC_MAX_MODES + j]; }
        }
        pr = prn;
        ph = ph + 1 == spt ? 0 : ph + 1;
        if (ph == 0) {
          if (g == 0 && in_slot) { tsx[ti] = (float)x; tsy[ti] = (float)y; }
          ti = (ti + 1) & (TS - 1);
          tcount++;
        }
        dc = dc + 1 == decim ? 0 : dc + 1;
        if (dc == 0) {
          hval = make_float2((float)(x * 1e6), (float)(y * 1e6));
          hidx = hpos;
          hpend = true;
          hpos = (hpos + 1) & (HIST - 1);
          hcount++;
        }
#pragma unroll
        for (int p = 0; p < MAXP; p++) { cur[p] = nxt[p]; csc[p] = nsc[p]; cbin[p] = nbin[p]; rp[p] = rpn[p]; }
        wave_sync();
      }
      if (hpend && g == 0 && in_slot) hist[hidx] = hval;   // the last step's sample

      // ---- chatter index per slot (src/cutting.js:299-318)
      if (in_slot && g == 0) TL.eng[gslot] = my_eng;
      if (g == 0 && in_slot) {
        double index = 0, amp = 0, once = 0;
        const int n = tcount < 24 ? tcount : 24;
        if (n >= 8) {
          double mx = 0, my = 0, v = 0;
          for (int k = 0; k < n; k++) { const int j = (ti - 1 - k) & (TS - 1); mx += tsx[j]; my += tsy[j]; }
          mx /= n; my /= n;
          for (int k = 0; k < n; k++) {
            const int j = (ti - 1 - k) & (TS - 1);
            const double dx = tsx[j] - mx, dy = tsy[j] - my;
            v += dx * dx + dy * dy;
          }
          once = __builtin_sqrt(v / n) * 1e6;
          const double r = mcm::round_js((double)n * (double)spt / (double)decim);
          const int m = (double)hcount < r ? hcount : (int)r;
          double ax = 0, ay = 0, w = 0;
          for (int k = 0; k < m; k++) { const float2 h = hist[(hpos - 1 - k) & (HIST - 1)]; ax += h.x; ay += h.y; }
          const double mm2 = m ? (double)m : 1.0;
          ax /= mm2; ay /= mm2;
          for (int k = 0; k < m; k++) {
            const float2 h = hist[(hpos - 1 - k) & (HIST - 1)];
            const double dx = h.x - ax, dy = h.y - ay;
            w += dx * dx + dy * dy;
          }
          amp = __builtin_sqrt(w / mm2);
          index = once / (amp + 0.05);
        }
        TL.idx[gslot] = index; TL.amp[gslot] = amp; TL.once[gslot] = once; TL.nz[gslot] = nz;
      }
      task_sync<WW>();
      x_trials += nnode;

      // ---- walk the tree with the verdicts (wave-uniform)
      auto take = [&](int j) {
        p_trials++;
        p_steps += steps;
        p_visits += (long long)N * TL.nz[j] * steps;
        p_eng += TL.eng[j];
        last_index = TL.idx[j]; last_amp = TL.amp[j]; last_once = TL.once[j]; last_nz = TL.nz[j];
        last_chat = TL.idx[j] > MC_CHATTER_IDX;
      };
      if (amax_task) {
        take(0);
        if (lead && P.cstate)
          __hip_atomic_store(P.cstate + cell, last_chat ? 1 : 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
      // tree nodes evaluated this round: follow the verdicts from the root while the
      // serial search's next node was among them
      int j = ntree > 0 ? 0 : -1;
      while (j >= 0) {
        if (!(it < iters && hi - lo > tol)) break;
        take(tfirst + j);
        const double m = TL.node_ap[tfirst + j];
        int nx;
        if (last_chat) { hi = m; nx = TL.node_kid[tfirst + j][0]; } else { lo = m; nx = TL.node_kid[tfirst + j][1]; }
        it++;
        j = nx >= 0 && nx < ntree ? nx : -1;
      }
      task_sync<WW>();  // the next round's plan overwrites the shared block
    }

    if (lead) {
      ResultDev& R = outp[task];
      R.alim = __builtin_sqrt(lo * hi);
      R.lo = lo; R.hi = hi;
      R.index = last_index; R.amp = last_amp; R.once = last_once;
      R.open = 0; R.trials = (int)p_trials; R.chatter = last_chat; R.nz = last_nz;
      R.fell_back = tab_overflow ? 2 : (int)failed; R.pad = (int)x_trials;
      R.steps = p_steps; R.visits = p_visits; R.engaged = p_eng;
      R.t_start = t_start; R.t_end = wall_clock64();
    }
    task_sync<WW>();  // the next task index overwrites the shared block
  }
}

// Persistent workgroups of four waves: the wide tasks first (a whole workgroup each),
// then one task per wave until the queue is empty.
__global__ __launch_bounds__(256) void cell_search(LaunchDev P) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned char* const wblk = lds + (size_t)w * per_wave_bytes();
  TaskLds& TLw = *reinterpret_cast<TaskLds*>(wblk);
  WaveLds<8>& W = *reinterpret_cast<WaveLds<8>*>(wblk + task_bytes());
  unsigned char* const area = wblk + task_bytes() + wave_bytes<8>();
  TaskLds& TL0 = *reinterpret_cast<TaskLds*>(lds);   // wave 0's task block: shared in wide tasks
  const int64_t wid = (int64_t)blockIdx.x * 4 + w;   // this wave's delay lines, rings, table
  float* const ring_base = P.ring + wid * (int64_t)8 * P.ring_cap;
  float2* const hist_base = P.hist + wid * (int64_t)8 * HIST;
  double2* const gsc = P.geo + wid * P.geo_cap;
  unsigned char* const gbn = P.gbin + wid * P.geo_cap;
  while (true) {
    if (threadIdx.x == 0) TL0.task = atomicAdd(P.wqueue, 1);
    __syncthreads();
    const int task = uni(TL0.task);
    __syncthreads();  // every wave has the index before wave 0 may reuse its block
    if (task >= P.nwide) break;
    run_task<4, 2>(P, P.wtasks, P.wout, task, TL0, W, area, w, lane, ring_base, hist_base, gsc, gbn);
  }
  while (true) {
    if (lane == 0) TLw.task = atomicAdd(P.queue, 1);
    wave_sync();
    const int task = uni(TLw.task);
    if (task >= P.ntasks) break;
    run_task<1, 8>(P, P.tasks, P.out, task, TLw, W, area, 0, lane, ring_base, hist_base, gsc, gbn);
  }
}

__global__ void reset_queue(int32_t* q, int n) {
  if ((int)threadIdx.x < n) q[threadIdx.x] = 0;
}
__global__ void clear_state(int32_t* s, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = 0;
}
int mcd_launch_clear_state(int32_t* s, int n, hipStream_t stream) {
  if (n <= 0) return 0;
  if (!s) return -2;
  hipLaunchKernelGGL(clear_state, dim3((n + 255) / 256), dim3(256), 0, stream, s, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
int mcd_launch_reset_queue(int32_t* q, int n, hipStream_t stream) {
  hipLaunchKernelGGL(reset_queue, dim3(1), dim3(64), 0, stream, q, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// blocks = persistent workgroups of four waves
int mcd_launch_search(const LaunchDev& P, int blocks, hipStream_t stream) {
  const size_t lds = (size_t)mcd_search_lds_bytes(P.nz_cap);
  hipLaunchKernelGGL(cell_search, dim3(blocks), dim3(256), lds, stream, P);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
