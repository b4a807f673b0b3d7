// lobes.hip — zero-order (Altintas-Budak) stability lobes of src/sld.js on the device:
// the lobe diagram of one parameter set, Simulator.lobes' composite curve over many speed
// nodes, and the analytic limit of many cells at their own speeds.
//
// stabilityLobes (src/sld.js:23-61) samples the chatter frequency over [0.8 fmin, 1.7 fmax]
// (NS = 400 * min(4, Mx + My - 1) samples), solves the 2 x 2 eigenvalue problem
// a0 L^2 + a1 L + 1 = 0 of the oriented transfer function at each sample, keeps the smaller
// positive depth, and maps each kept sample to one spindle speed per lobe number j; a lobe
// is its samples inside the speed window, sorted by speed (Array.prototype.sort, stable).
// alimAt (src/sld.js:64-76) takes the lowest lobe covering a speed, interpolated linearly.
//
// The kernels:
//   lobe_boundary  one thread per (job, sample): the boundary samples (wc, eps = pi - 2 psi,
//                  alim; alim = 0 marks a sample without a positive depth) into HBM;
//   lobe_emit,     one 256-thread workgroup per (job, lobe j): compacts the samples whose
//   lobe_read
//                  speed falls in the window into LDS, bitonic-sorts them by (speed, sample)
//                  — the order of a stable sort by speed — and then either writes the lobe
//                  (the diagram) or reads it at the job's speeds: a binary search finds the
//                  first segment that can cover S, the covering segments follow in order, and
//                  the lobe's minimum goes into the speed's running minimum by a 64-bit
//                  atomic min.
// The atomic min equals the reference's sequential "if (v < best) best = v" over lobes and
// segments: every candidate is +0, positive or NaN (a > 0, b > 0, t in [0, 1] make
// a + (b - a) t >= +0 after rounding, and x + (-x) is +0), NaN never wins the comparison,
// and the bit patterns of non-negative doubles order like the values, so the minimum is
// one bit pattern whatever the order.  The arithmetic follows the JavaScript operation by
// operation (fdlibm sin / cos / atan, V8's hypot), so the results equal the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_host.h"
#include "mc_math.h"

namespace {

constexpr double JS_PI = 3.141592653589793;
constexpr int NS_MAX = 1600;      // 400 * 4 samples
constexpr int SORT_CAP = 2048;    // power of two >= NS_MAX
constexpr int WG = 256;
constexpr uint64_t SENTINEL = ~0ull;

// one boundary (a parameter set) and what is read from its lobes
struct LobeJob {
  mc_mode mx[MC_MAX_MODES], my[MC_MAX_MODES];
  int32_t nmx, nmy, N, NS;
  double Ktc, a0, a3, det;   // axx, ayy and det of the directional coefficients
  double f0, f1;             // chatter-frequency range
  double lo, hi;             // lobe speed window
  int64_t q0, nq;            // the job's speeds: queries [q0, q0 + nq)
};

struct LobeTask { int32_t job, j; };

// receptance sum over modes (src/frf.js:14-21), k in N/mm
__device__ inline void frf(const mc_mode* m, int n, double f, double& re, double& im) {
  re = 0;
  im = 0;
  for (int i = 0; i < n; i++) {
    const double r = f / m[i].fn, a = 1 - r * r, b = 2 * m[i].zeta * r, d = m[i].k * (a * a + b * b);
    re += a / d;
    im -= b / d;
  }
}

__device__ inline double js_sign(double v) {
  if (v != v) return __builtin_nan("");
  if (v > 0) return 1;
  if (v < 0) return -1;
  return v;
}

// boundary samples (src/sld.js:27-52): one thread per (sample, job)
__global__ __launch_bounds__(WG) void lobe_boundary(const LobeJob* __restrict__ jobs, double* __restrict__ wcO,
                                                    double* __restrict__ epsO, double* __restrict__ alimO) {
  const LobeJob& a = jobs[blockIdx.y];
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= a.NS) return;
  const int NS = a.NS;
  const double fc = a.f0 + (a.f1 - a.f0) * s / (NS - 1), wc = 2 * JS_PI * fc;
  double pxr, pxi, pyr, pyi;
  frf(a.mx, a.nmx, fc, pxr, pxi);
  frf(a.my, a.nmy, fc, pyr, pyi);
  const double det = a.det;
  const double ppr = pxr * pyr - pxi * pyi, ppi = pxr * pyi + pxi * pyr;
  const double a0r = ppr * det - ppi * 0, a0i = ppr * 0 + ppi * det;
  const double a1r = a.a0 * pxr + a.a3 * pyr, a1i = a.a0 * pxi + a.a3 * pyi;
  const double sqr = (a1r * a1r - a1i * a1i) - 4 * a0r, sqi = (a1r * a1i + a1i * a1r) - 4 * a0i;
  // complex square root of sq (principal branch, as the reference writes it)
  const double r = mcm::hypot_js(sqr, sqi);
  const double sgn = js_sign((sqi != 0 && sqi == sqi) ? sqi : 1);
  const double t = (r - sqr) / 2;
  const double dr = __builtin_sqrt((r + sqr) / 2), di = sgn * __builtin_sqrt(t > 0 ? t : 0);
  bool has = false;
  double bestA = 0, bestK = 0;
  for (int g = 0; g < 2; g++) {
    const double sg = g ? -1 : 1;
    const double nr = -(a1r + sg * dr), ni = -(a1i + sg * di);
    const double der = 2 * a0r, dei = 2 * a0i;
    const double d = der * der + dei * dei;
    const double Lr = (nr * der + ni * dei) / d, Li = (ni * der - nr * dei) / d;
    if (Lr >= 0) continue;
    const double kap = Li / Lr;
    const double alim = -2 * JS_PI * Lr * (1 + kap * kap) / (a.N * a.Ktc);
    if (alim > 0 && (!has || alim < bestA)) { bestA = alim; bestK = kap; has = true; }
  }
  const size_t o = (size_t)blockIdx.y * NS_MAX + s;
  wcO[o] = wc;
  epsO[o] = has ? JS_PI - 2 * mcm::atan_js(bestK) : 0.0;
  alimO[o] = has ? bestA : 0.0;
}

struct LobeLds {
  uint64_t key[SORT_CAP];   // speed bits (positive doubles order like their bits)
  uint16_t idx[SORT_CAP];   // sample number: the stable sort's tie break
  int n;
};

// lobe j of the job in LDS, sorted by (speed, sample); returns its length
__device__ int build_lobe(LobeLds& L, const LobeJob& a, int j, const double* wc, const double* eps,
                          const double* alim) {
  const int tid = threadIdx.x;
  if (tid == 0) L.n = 0;
  __syncthreads();
  for (int s = tid; s < a.NS; s += WG) {
    if (!(alim[s] > 0)) continue;
    const double T = (eps[s] + 2 * j * JS_PI) / wc[s];
    const double r = 60 / (a.N * T);
    if (r >= a.lo && r <= a.hi) {
      const int slot = atomicAdd(&L.n, 1);
      L.key[slot] = __builtin_bit_cast(uint64_t, r);
      L.idx[slot] = (uint16_t)s;
    }
  }
  __syncthreads();
  const int n = L.n;
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = n + tid; i < P; i += WG) { L.key[i] = SENTINEL; L.idx[i] = 0xffff; }
  __syncthreads();
  // bitonic sort, ascending, every thread on a pair of each stage
  for (int k = 2; k <= P; k <<= 1) {
    for (int h = k >> 1; h > 0; h >>= 1) {
      for (int p = tid; p < (P >> 1); p += WG) {
        const int i = ((p & ~(h - 1)) << 1) | (p & (h - 1)), q = i | h;
        const uint64_t ki = L.key[i], kq = L.key[q];
        const uint16_t ii = L.idx[i], iq = L.idx[q];
        const bool gt = ki > kq || (ki == kq && ii > iq);
        if (gt == ((i & k) == 0)) { L.key[i] = kq; L.key[q] = ki; L.idx[i] = iq; L.idx[q] = ii; }
      }
      __syncthreads();
    }
  }
  return n;
}

// the lobe diagram: lobe j of job 0 into out[task * NS_MAX ...], its length into len[task]
__global__ __launch_bounds__(WG) void lobe_emit(const LobeJob* __restrict__ jobs, const LobeTask* __restrict__ tasks,
                                                const double* __restrict__ wc, const double* __restrict__ eps,
                                                const double* __restrict__ alim, double* __restrict__ rpmO,
                                                double* __restrict__ alimO, int* __restrict__ len) {
  __shared__ LobeLds L;
  const LobeTask tk = tasks[blockIdx.x];
  const size_t so = (size_t)tk.job * NS_MAX;
  const int n = build_lobe(L, jobs[tk.job], tk.j, wc + so, eps + so, alim + so);
  const size_t o = (size_t)blockIdx.x * NS_MAX;
  for (int i = threadIdx.x; i < n; i += WG) {
    rpmO[o + i] = __builtin_bit_cast(double, L.key[i]);
    alimO[o + i] = alim[so + L.idx[i]];
  }
  if (threadIdx.x == 0) len[blockIdx.x] = n;
}

// lobe j of the job read at the job's speeds; the lobe's lowest covering value goes into
// best[q] (bits of a non-negative double, +inf where nothing covers yet)
__global__ __launch_bounds__(WG) void lobe_read(const LobeJob* __restrict__ jobs, const LobeTask* __restrict__ tasks,
                                                const double* __restrict__ wc, const double* __restrict__ eps,
                                                const double* __restrict__ alim, const double* __restrict__ S,
                                                unsigned long long* __restrict__ best) {
  __shared__ LobeLds L;
  const LobeTask tk = tasks[blockIdx.x];
  const LobeJob& a = jobs[tk.job];
  const size_t so = (size_t)tk.job * NS_MAX;
  const int n = build_lobe(L, a, tk.j, wc + so, eps + so, alim + so);
  if (n < 2) return;
  const double* al = alim + so;
  for (int64_t q = a.q0 + threadIdx.x; q < a.q0 + a.nq; q += WG) {
    const double s = S[q];
    // lb = the number of speeds below s: segments i-1, i with i >= lb may cover s
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (__builtin_bit_cast(double, L.key[mid]) < s) lo = mid + 1; else hi = mid;
    }
    double b = __builtin_inf();
    for (int i = lo > 1 ? lo : 1; i < n; i++) {
      const double ar = __builtin_bit_cast(double, L.key[i - 1]), br = __builtin_bit_cast(double, L.key[i]);
      if (!(ar <= s)) break;   // later segments start above s
      if (s >= ar && s <= br) {
        const double d = br - ar;
        const double t = (s - ar) / (d != 0 ? d : 1);
        const double a0 = al[L.idx[i - 1]], a1 = al[L.idx[i]];
        const double v = a0 + (a1 - a0) * t;
        if (v < b) b = v;
      }
    }
    if (b < __builtin_inf()) atomicMin(best + q, __builtin_bit_cast(unsigned long long, b));
  }
}

__global__ void fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

struct Buf {
  void* p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return MC_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, bytes < 256 ? 256 : bytes) != hipSuccess) return mch::fail(MC_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    n = bytes < 256 ? 256 : bytes;
    return MC_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// device buffers of the lobe entry points (grow-only, guarded by the library's call lock)
struct LobeBufs {
  Buf jobs, tasks, wc, eps, alim, S, best, rpmO, alimO, len;
};
LobeBufs& bufs() {
  static LobeBufs* b = new LobeBufs();   // lives until the process ends
  return *b;
}

int check_args(const mc_lobe_args* a) {
  if (!a) return mch::fail(MC_ERR_ARG, "null lobe arguments");
  if (a->N < 1 || a->nmx < 1 || a->nmx > MC_MAX_MODES || a->nmy < 1 || a->nmy > MC_MAX_MODES)
    return mch::fail(MC_ERR_ARG, "need N >= 1 and 1..%d modes per axis", MC_MAX_MODES);
  if (a->lobes < 0 || a->lobes > 100000) return mch::fail(MC_ERR_ARG, "bad lobe count");
  return MC_OK;
}

// the boundary of a parameter set and its lobe speed window [lo, hi]
LobeJob make_job(int N, double Ktc, double Krc, double phiSt, double phiEx, int nmx, const mc_mode* mx, int nmy,
                 const mc_mode* my, double lo, double hi) {
  LobeJob j;
  std::memset(&j, 0, sizeof j);
  double al[4];
  mc_directional_coeffs(Krc / Ktc, phiSt, phiEx, al);
  j.a0 = al[0];
  j.a3 = al[3];
  j.det = al[0] * al[3] - al[1] * al[2];
  double fmin = INFINITY, fmax = -INFINITY;
  for (int i = 0; i < nmx; i++) { fmin = std::min(fmin, mx[i].fn); fmax = std::max(fmax, mx[i].fn); j.mx[i] = mx[i]; }
  for (int i = 0; i < nmy; i++) { fmin = std::min(fmin, my[i].fn); fmax = std::max(fmax, my[i].fn); j.my[i] = my[i]; }
  j.f0 = 0.8 * fmin;
  j.f1 = 1.7 * fmax;
  j.NS = 400 * std::min(4, nmx + nmy - 1);
  j.nmx = nmx;
  j.nmy = nmy;
  j.N = N;
  j.Ktc = Ktc;
  j.lo = lo;
  j.hi = hi;
  return j;
}

LobeJob job_of(const mc_lobe_args& a) {
  return make_job(a.N, a.Ktc, a.Krc, a.phiSt, a.phiEx, a.nmx, a.mx, a.nmy, a.my, a.rpmMin * 0.8, a.rpmMax * 1.2);
}

// uploads the jobs and computes their boundaries
int run_boundaries(const std::vector<LobeJob>& jobs, hipStream_t s) {
  LobeBufs& B = bufs();
  const size_t nj = jobs.size(), ns = nj * NS_MAX * sizeof(double);
  int st;
  if ((st = B.jobs.ensure(nj * sizeof(LobeJob))) || (st = B.wc.ensure(ns)) || (st = B.eps.ensure(ns)) ||
      (st = B.alim.ensure(ns)))
    return st;
  MCH_HIPCHK(hipMemcpyAsync(B.jobs.p, jobs.data(), nj * sizeof(LobeJob), hipMemcpyHostToDevice, s));
  for (size_t j0 = 0; j0 < nj; j0 += 65535) {
    const int nb = (int)std::min<size_t>(65535, nj - j0);
    hipLaunchKernelGGL(lobe_boundary, dim3((NS_MAX + WG - 1) / WG, nb), dim3(WG), 0, s, B.jobs.as<LobeJob>() + j0,
                       B.wc.as<double>() + j0 * NS_MAX, B.eps.as<double>() + j0 * NS_MAX,
                       B.alim.as<double>() + j0 * NS_MAX);
  }
  MCH_HIPCHK(hipGetLastError());
  return MC_OK;
}

// every lobe of every job read at the jobs' speeds S (host, nq values): best[q] = the lowest
// covering value (+inf if none)
int read_lobes(const std::vector<LobeJob>& jobs, const std::vector<int>& lobes, const std::vector<double>& S,
               std::vector<double>& best) {
  hipStream_t s = mch::stream();
  LobeBufs& B = bufs();
  int st = run_boundaries(jobs, s);
  if (st) return st;
  std::vector<LobeTask> tasks;
  for (size_t i = 0; i < jobs.size(); i++)
    for (int j = 0; j < lobes[i]; j++) tasks.push_back({(int32_t)i, j});
  const size_t nq = S.size();
  if ((st = B.S.ensure(nq * 8)) || (st = B.best.ensure(nq * 8)) || (st = B.tasks.ensure(tasks.size() * sizeof(LobeTask))))
    return st;
  best.assign(nq, INFINITY);
  if (nq) {
    MCH_HIPCHK(hipMemcpyAsync(B.S.p, S.data(), nq * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_u64, dim3((unsigned)std::min<size_t>(1024, (nq + WG - 1) / WG)), dim3(WG), 0, s,
                       B.best.as<unsigned long long>(), (int64_t)nq, __builtin_bit_cast(unsigned long long, __builtin_inf()));
  }
  if (nq && !tasks.empty()) {
    MCH_HIPCHK(hipMemcpyAsync(B.tasks.p, tasks.data(), tasks.size() * sizeof(LobeTask), hipMemcpyHostToDevice, s));
    // at most 2^24 workgroups of 256 per launch (2^32 work-items)
    for (size_t t0 = 0; t0 < tasks.size(); t0 += 1u << 24) {
      const unsigned nb = (unsigned)std::min<size_t>(1u << 24, tasks.size() - t0);
      hipLaunchKernelGGL(lobe_read, dim3(nb), dim3(WG), 0, s, B.jobs.as<LobeJob>(), B.tasks.as<LobeTask>() + t0,
                         B.wc.as<double>(), B.eps.as<double>(), B.alim.as<double>(), B.S.as<double>(),
                         B.best.as<unsigned long long>());
    }
  }
  MCH_HIPCHK(hipGetLastError());
  if (nq) MCH_HIPCHK(hipMemcpyAsync(best.data(), B.best.p, nq * 8, hipMemcpyDeviceToHost, s));
  MCH_HIPCHK(hipStreamSynchronize(s));
  return MC_OK;
}

// speeds a composite node reads: rpmMin, rpmMin + step, ... below min(hi, rpmMax)
int64_t node_points(const mc_lobe_args& a, double step, double hi) {
  const double top = std::min(hi, a.rpmMax);
  int64_t n = 0;
  for (double S = a.rpmMin; S < top; S += step) n++;
  return n;
}

}  // namespace

extern "C" {

int mc_directional_coeffs(double Kr, double phiSt, double phiEx, double* out) {
  if (!out) return mch::fail(MC_ERR_ARG, "null output");
  double a[4], b[4];
  for (int t = 0; t < 2; t++) {
    double* o = t ? b : a;
    const double p = t ? phiSt : phiEx;
    o[0] = 0.5 * (mcm::cos_js(2 * p) - 2 * Kr * p + Kr * mcm::sin_js(2 * p));
    o[1] = 0.5 * (-mcm::sin_js(2 * p) - 2 * p + Kr * mcm::cos_js(2 * p));
    o[2] = 0.5 * (-mcm::sin_js(2 * p) + 2 * p + Kr * mcm::cos_js(2 * p));
    o[3] = 0.5 * (-mcm::cos_js(2 * p) - 2 * Kr * p - Kr * mcm::sin_js(2 * p));
  }
  for (int t = 0; t < 4; t++) out[t] = a[t] - b[t];
  return MC_OK;
}

int mc_stability_lobes(const mc_lobe_args* a, double* rpm, double* alim, int max_pts, int* lobe_len, int* n_out) {
  int st = check_args(a);
  if (st) return st;
  if (max_pts < 0 || (max_pts > 0 && (!rpm || !alim)) || (a->lobes > 0 && !lobe_len) || !n_out)
    return mch::fail(MC_ERR_ARG, "bad output arrays");
  if ((st = mch::init())) return st;
  std::lock_guard<std::mutex> lk(mch::lock());
  hipStream_t s = mch::stream();
  LobeBufs& B = bufs();
  std::vector<LobeJob> jobs{job_of(*a)};
  if ((st = run_boundaries(jobs, s))) return st;
  // lobes in chunks of CH, each lobe written sorted into its own NS_MAX stretch
  constexpr int CH = 1024;
  const int nl = a->lobes;
  if ((st = B.rpmO.ensure((size_t)CH * NS_MAX * 8)) || (st = B.alimO.ensure((size_t)CH * NS_MAX * 8)) ||
      (st = B.len.ensure(CH * 4)) || (st = B.tasks.ensure(CH * sizeof(LobeTask))))
    return st;
  std::vector<double> hr((size_t)CH * NS_MAX), ha((size_t)CH * NS_MAX);
  std::vector<int> hl(CH);
  std::vector<LobeTask> tasks(CH);
  int total = 0, nout = 0;
  for (int j0 = 0; j0 < nl; j0 += CH) {
    const int nb = std::min(CH, nl - j0);
    for (int t = 0; t < nb; t++) tasks[t] = {0, j0 + t};
    MCH_HIPCHK(hipMemcpyAsync(B.tasks.p, tasks.data(), nb * sizeof(LobeTask), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(lobe_emit, dim3(nb), dim3(WG), 0, s, B.jobs.as<LobeJob>(), B.tasks.as<LobeTask>(),
                       B.wc.as<double>(), B.eps.as<double>(), B.alim.as<double>(), B.rpmO.as<double>(),
                       B.alimO.as<double>(), B.len.as<int>());
    MCH_HIPCHK(hipGetLastError());
    MCH_HIPCHK(hipMemcpyAsync(hl.data(), B.len.p, nb * 4, hipMemcpyDeviceToHost, s));
    MCH_HIPCHK(hipMemcpyAsync(hr.data(), B.rpmO.p, (size_t)nb * NS_MAX * 8, hipMemcpyDeviceToHost, s));
    MCH_HIPCHK(hipMemcpyAsync(ha.data(), B.alimO.p, (size_t)nb * NS_MAX * 8, hipMemcpyDeviceToHost, s));
    MCH_HIPCHK(hipStreamSynchronize(s));
    // lobes with no point are left out; a lobe's full length is recorded even when max_pts
    // cuts its points short
    for (int t = 0; t < nb; t++) {
      if (hl[t] == 0) continue;
      for (int i = 0; i < hl[t] && total < max_pts; i++, total++) {
        rpm[total] = hr[(size_t)t * NS_MAX + i];
        alim[total] = ha[(size_t)t * NS_MAX + i];
      }
      lobe_len[nout++] = hl[t];
    }
  }
  for (int j = nout; j < nl; j++) lobe_len[j] = 0;
  *n_out = total;
  return MC_OK;
}

int mc_alim_at(const double* rpm, const double* alim, const int* lobe_len, int nlobes, double S, double* out) {
  if (!out || nlobes < 0 || (nlobes > 0 && (!lobe_len || !rpm || !alim))) return mch::fail(MC_ERR_ARG, "bad arguments");
  double best = INFINITY;
  int off = 0;
  for (int j = 0; j < nlobes; j++) {
    for (int i = 1; i < lobe_len[j]; i++) {
      const double ar = rpm[off + i - 1], br = rpm[off + i];
      if (S >= ar && S <= br) {
        const double d = br - ar;
        const double t = (S - ar) / (d != 0 ? d : 1);
        const double v = alim[off + i - 1] + (alim[off + i] - alim[off + i - 1]) * t;
        if (v < best) best = v;
      }
    }
    off += lobe_len[j];
  }
  *out = best;
  return MC_OK;
}

// Simulator.lobes' composite curve (src/sim.js:270-281): node i's diagram read at its speeds
// rpmMin, rpmMin + step, ... (accumulated as the reference does) — every (node, lobe) is one
// workgroup of lobe_read
int mc_composite_lobes(const mc_lobe_args* nodes, int n, double step, double hi, double* rpm, double* alim,
                       int max_pts, int* n_out) {
  if (n < 0 || (n > 0 && !nodes) || !n_out) return mch::fail(MC_ERR_ARG, "bad arguments");
  if (!(step > 0) || !std::isfinite(step) || std::isnan(hi)) return mch::fail(MC_ERR_ARG, "need a finite step > 0");
  std::vector<int64_t> off(n + 1, 0);
  for (int i = 0; i < n; i++) {
    int st = check_args(&nodes[i]);
    if (st) return st;
    const mc_lobe_args& a = nodes[i];
    if (!std::isfinite(a.rpmMin) || !std::isfinite(a.rpmMax)) return mch::fail(MC_ERR_ARG, "node %d: bad speed window", i);
    if (std::min(hi, a.rpmMax) - a.rpmMin > step * 1e7) return mch::fail(MC_ERR_ARG, "node %d: over 1e7 speeds", i);
    off[i + 1] = off[i] + node_points(a, step, hi);
    if (off[i + 1] > 0x7fffffff) return mch::fail(MC_ERR_ARG, "too many points");
  }
  *n_out = (int)off[n];
  if (!rpm && !alim && max_pts == 0) return MC_OK;  // size query
  if (!rpm || !alim || max_pts < off[n]) return mch::fail(MC_ERR_ARG, "the curve needs %lld points", (long long)off[n]);
  if (off[n] == 0) return MC_OK;
  int st = mch::init();
  if (st) return st;
  std::lock_guard<std::mutex> lk(mch::lock());
  std::vector<LobeJob> jobs(n);
  std::vector<int> lobes(n);
  std::vector<double> S((size_t)off[n]), best;
  for (int i = 0; i < n; i++) {
    jobs[i] = job_of(nodes[i]);
    jobs[i].q0 = off[i];
    jobs[i].nq = off[i + 1] - off[i];
    lobes[i] = nodes[i].lobes;
    double v = nodes[i].rpmMin;
    for (int64_t q = off[i]; q < off[i + 1]; q++, v += step) S[q] = v;
  }
  if ((st = read_lobes(jobs, lobes, S, best))) return st;
  for (int64_t q = 0; q < off[n]; q++) {
    rpm[q] = S[q];
    alim[q] = best[q] == INFINITY ? 1e3 : best[q];
  }
  return MC_OK;
}

// alimAt of every cell at its own speed: cells that differ only in speed share one boundary
// and one set of lobes (enough lobe numbers to reach their lowest speed)
int mc_analytic_limits(const mc_cell* cells, int n, double* out) {
  if (n < 0 || (n > 0 && (!cells || !out))) return mch::fail(MC_ERR_ARG, "bad arguments");
  std::map<std::vector<double>, std::vector<int>> groups;
  for (int i = 0; i < n; i++) {
    const mc_cell& c = cells[i];
    if (c.tool.N < 1 || c.nmx < 1 || c.nmx > MC_MAX_MODES || c.nmy < 1 || c.nmy > MC_MAX_MODES)
      return mch::fail(MC_ERR_ARG, "cell %d: bad flutes or modes", i);
    std::vector<double> key = {(double)c.tool.N, c.mat.Ktc, c.mat.Krc, c.phiSt, c.phiEx, (double)c.nmx, (double)c.nmy};
    for (int j = 0; j < c.nmx; j++) { key.push_back(c.mx[j].k); key.push_back(c.mx[j].fn); key.push_back(c.mx[j].zeta); }
    for (int j = 0; j < c.nmy; j++) { key.push_back(c.my[j].k); key.push_back(c.my[j].fn); key.push_back(c.my[j].zeta); }
    groups[key].push_back(i);
  }
  if (n == 0) return MC_OK;
  int st = mch::init();
  if (st) return st;
  std::lock_guard<std::mutex> lk(mch::lock());
  std::vector<LobeJob> jobs;
  std::vector<int> lobes, order;
  std::vector<double> S, best;
  for (auto& g : groups) {
    const std::vector<int>& idx = g.second;
    const mc_cell& c = cells[idx[0]];
    double smin = INFINITY, smax = 0, fmax = 0;
    for (int i : idx) { smin = std::min(smin, cells[i].S); smax = std::max(smax, cells[i].S); }
    for (int j = 0; j < c.nmx; j++) fmax = std::max(fmax, c.mx[j].fn);
    for (int j = 0; j < c.nmy; j++) fmax = std::max(fmax, c.my[j].fn);
    LobeJob jb = make_job(c.tool.N, c.mat.Ktc, c.mat.Krc, c.phiSt, c.phiEx, c.nmx, c.mx, c.nmy, c.my, smin * 0.8,
                          smax * 1.2);
    jb.q0 = (int64_t)S.size();
    jb.nq = (int64_t)idx.size();
    jobs.push_back(jb);
    // lobe j maps chatter frequency fc to rpm <= 60 fc / (N j): beyond
    // 60 * 1.7 fmax / (N * 0.8 smin) no lobe reaches the lowest speed
    lobes.push_back((int)std::min(100000.0, std::ceil(60 * 1.7 * fmax / (c.tool.N * 0.8 * smin)) + 2));
    for (int i : idx) { S.push_back(cells[i].S); order.push_back(i); }
  }
  if ((st = read_lobes(jobs, lobes, S, best))) return st;
  for (size_t q = 0; q < order.size(); q++) out[order[q]] = best[q];
  return MC_OK;
}

}  // extern "C"
