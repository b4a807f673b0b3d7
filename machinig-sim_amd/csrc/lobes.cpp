// lobes.cpp — zero-order (Altintas-Budak) stability lobes of src/sld.js on the host,
// and the analytic limit of many cells at their own speeds.
//
// stabilityLobes samples the chatter frequency over [0.8 fmin, 1.7 fmax]
// (400 * min(4, Mx + My - 1) samples), solves the 2 x 2 eigenvalue problem
// a0 L^2 + a1 L + 1 = 0 of the oriented transfer function at each sample, keeps the
// smaller positive depth, and maps each kept sample to one spindle speed per lobe number;
// alimAt takes the lowest lobe covering a speed, interpolated linearly.  The arithmetic
// follows the JavaScript operation by operation (fdlibm sin / cos / atan / hypot), so the
// results equal the reference's.  These limits are cheap (no time-domain simulation) and
// serve as a physics input of the bracket surrogate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_host.h"
#include "mc_math.h"

namespace {

constexpr double JS_PI = 3.141592653589793;

// receptance sum over modes (src/frf.js:14-21), k in N/mm
void frf(const mc_mode* m, int n, double f, double& re, double& im) {
  re = 0;
  im = 0;
  for (int i = 0; i < n; i++) {
    const double r = f / m[i].fn, a = 1 - r * r, b = 2 * m[i].zeta * r, d = m[i].k * (a * a + b * b);
    re += a / d;
    im -= b / d;
  }
}

double js_sign(double v) {
  if (v != v) return NAN;
  if (v > 0) return 1;
  if (v < 0) return -1;
  return v;
}

struct Sample { double wc, alim, psi; };
struct LobePt { double rpm, alim; };

// samples of the stability boundary over chatter frequency (src/sld.js:27-52)
void boundary(const mc_lobe_args& a, std::vector<Sample>& pts) {
  double al[4];
  mc_directional_coeffs(a.Krc / a.Ktc, a.phiSt, a.phiEx, al);
  const double det = al[0] * al[3] - al[1] * al[2];
  double fmin = INFINITY, fmax = -INFINITY;
  for (int i = 0; i < a.nmx; i++) { fmin = std::min(fmin, a.mx[i].fn); fmax = std::max(fmax, a.mx[i].fn); }
  for (int i = 0; i < a.nmy; i++) { fmin = std::min(fmin, a.my[i].fn); fmax = std::max(fmax, a.my[i].fn); }
  const double f0 = 0.8 * fmin, f1 = 1.7 * fmax;
  const int NS = 400 * std::min(4, a.nmx + a.nmy - 1);
  pts.clear();
  for (int s = 0; s < NS; s++) {
    const double fc = f0 + (f1 - f0) * s / (NS - 1), wc = 2 * JS_PI * fc;
    double pxr, pxi, pyr, pyi;
    frf(a.mx, a.nmx, fc, pxr, pxi);
    frf(a.my, a.nmy, fc, pyr, pyi);
    const double ppr = pxr * pyr - pxi * pyi, ppi = pxr * pyi + pxi * pyr;
    const double a0r = ppr * det - ppi * 0, a0i = ppr * 0 + ppi * det;
    const double a1r = al[0] * pxr + al[3] * pyr, a1i = al[0] * pxi + al[3] * pyi;
    const double sqr = (a1r * a1r - a1i * a1i) - 4 * a0r, sqi = (a1r * a1i + a1i * a1r) - 4 * a0i;
    // complex square root of sq (principal branch, as the reference writes it)
    const double r = mcm::hypot_js(sqr, sqi);
    const double sgn = js_sign((sqi != 0 && sqi == sqi) ? sqi : 1);
    const double t = (r - sqr) / 2;
    const double dr = std::sqrt((r + sqr) / 2), di = sgn * std::sqrt(t > 0 ? t : 0);
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
    if (has) pts.push_back({wc, bestA, mcm::atan_js(bestK)});
  }
}

// lobe j of the boundary samples, sorted by speed (a stable insertion sort keeps the
// order of equal speeds, as Array.prototype.sort does)
void lobe(const mc_lobe_args& a, const std::vector<Sample>& pts, int j, double lo, double hi, std::vector<LobePt>& out) {
  out.clear();
  for (const Sample& p : pts) {
    const double eps = JS_PI - 2 * p.psi;
    const double T = (eps + 2 * j * JS_PI) / p.wc;
    const double r = 60 / (a.N * T);
    if (r >= lo && r <= hi) out.push_back({r, p.alim});
  }
  for (size_t i = 1; i < out.size(); i++) {
    const LobePt v = out[i];
    size_t k = i;
    while (k > 0 && out[k - 1].rpm > v.rpm) { out[k] = out[k - 1]; k--; }
    out[k] = v;
  }
}

// the lowest lobe covering S, linearly interpolated (alimAt's inner loops, src/sld.js:66-73)
void alim_on(const std::vector<LobePt>& L, double S, double& best) {
  for (size_t i = 1; i < L.size(); i++) {
    const double ar = L[i - 1].rpm, br = L[i].rpm;
    if (S >= ar && S <= br) {
      const double d = br - ar;
      const double t = (S - ar) / (d != 0 ? d : 1);
      const double v = L[i - 1].alim + (L[i].alim - L[i - 1].alim) * t;
      if (v < best) best = v;
    }
  }
}

// speeds a composite node reads: rpmMin, rpmMin + step, ... below min(hi, rpmMax)
int64_t node_points(const mc_lobe_args& a, double step, double hi) {
  const double top = std::min(hi, a.rpmMax);
  int64_t n = 0;
  for (double S = a.rpmMin; S < top; S += step) n++;
  return n;
}

int check_args(const mc_lobe_args* a) {
  if (!a) return mch::fail(MC_ERR_ARG, "null lobe arguments");
  if (a->N < 1 || a->nmx < 1 || a->nmx > MC_MAX_MODES || a->nmy < 1 || a->nmy > MC_MAX_MODES)
    return mch::fail(MC_ERR_ARG, "need N >= 1 and 1..%d modes per axis", MC_MAX_MODES);
  if (a->lobes < 0 || a->lobes > 100000) return mch::fail(MC_ERR_ARG, "bad lobe count");
  return MC_OK;
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
  std::vector<Sample> pts;
  boundary(*a, pts);
  std::vector<LobePt> L;
  int total = 0, nout = 0;
  for (int j = 0; j < a->lobes; j++) {
    lobe(*a, pts, j, a->rpmMin * 0.8, a->rpmMax * 1.2, L);
    if (L.empty()) continue;
    for (size_t i = 0; i < L.size() && total < max_pts; i++, total++) { rpm[total] = L[i].rpm; alim[total] = L[i].alim; }
    lobe_len[nout++] = (int)L.size();
  }
  for (int j = nout; j < a->lobes; j++) lobe_len[j] = 0;
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

// Simulator.lobes' composite curve (src/sim.js:270-281): every node's diagram is built and
// read independently, so the nodes (of one curve or of many curves) run on host threads and
// each writes its own stretch of the output
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
  const int nthreads = std::max(1, std::min<int>(n, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) {
    pool.emplace_back([&, t]() {
      std::vector<Sample> pts;
      std::vector<LobePt> L;
      std::vector<double> best;
      for (int i = t; i < n; i += nthreads) {
        const mc_lobe_args& a = nodes[i];
        const int64_t m = off[i + 1] - off[i];
        best.assign((size_t)m, INFINITY);
        boundary(a, pts);
        for (int j = 0; j < a.lobes; j++) {
          lobe(a, pts, j, a.rpmMin * 0.8, a.rpmMax * 1.2, L);
          double S = a.rpmMin;
          for (int64_t q = 0; q < m; q++, S += step) alim_on(L, S, best[q]);
        }
        double S = a.rpmMin;
        for (int64_t q = 0; q < m; q++, S += step) {
          rpm[off[i] + q] = S;
          alim[off[i] + q] = best[q] == INFINITY ? 1e3 : best[q];
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  return MC_OK;
}

// alimAt of every cell at its own speed: cells that differ only in speed share one set of
// lobes (enough lobe numbers to reach their lowest speed); groups run on host threads
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
  std::vector<const std::vector<int>*> list;
  for (auto& g : groups) list.push_back(&g.second);
  const int nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) {
    pool.emplace_back([&, t]() {
      std::vector<Sample> pts;
      std::vector<LobePt> L;
      for (size_t gi = t; gi < list.size(); gi += nthreads) {
        const std::vector<int>& idx = *list[gi];
        const mc_cell& c = cells[idx[0]];
        mc_lobe_args a;
        std::memset(&a, 0, sizeof a);
        a.N = c.tool.N; a.Ktc = c.mat.Ktc; a.Krc = c.mat.Krc; a.phiSt = c.phiSt; a.phiEx = c.phiEx;
        a.nmx = c.nmx; a.nmy = c.nmy;
        for (int j = 0; j < c.nmx; j++) a.mx[j] = c.mx[j];
        for (int j = 0; j < c.nmy; j++) a.my[j] = c.my[j];
        double smin = INFINITY, smax = 0, fmax = 0;
        for (int i : idx) { smin = std::min(smin, cells[i].S); smax = std::max(smax, cells[i].S); }
        for (int j = 0; j < c.nmx; j++) fmax = std::max(fmax, c.mx[j].fn);
        for (int j = 0; j < c.nmy; j++) fmax = std::max(fmax, c.my[j].fn);
        boundary(a, pts);
        // lobe j maps chatter frequency fc to rpm <= 60 fc / (N j): beyond
        // 60 * 1.7 fmax / (N * 0.8 smin) no lobe reaches the lowest speed
        const int lobes = (int)std::min(100000.0, std::ceil(60 * 1.7 * fmax / (c.tool.N * 0.8 * smin)) + 2);
        std::vector<double> best(idx.size(), INFINITY);
        for (int j = 0; j < lobes; j++) {
          lobe(a, pts, j, smin * 0.8, smax * 1.2, L);
          for (size_t q = 0; q < idx.size(); q++) {
            const double S = cells[idx[q]].S;
            for (size_t i = 1; i < L.size(); i++) {
              const double ar = L[i - 1].rpm, br = L[i].rpm;
              if (S >= ar && S <= br) {
                const double d = br - ar;
                const double tt = (S - ar) / (d != 0 ? d : 1);
                const double v = L[i - 1].alim + (L[i].alim - L[i - 1].alim) * tt;
                if (v < best[q]) best[q] = v;
              }
            }
          }
        }
        for (size_t q = 0; q < idx.size(); q++) out[idx[q]] = best[q];
      }
    });
  }
  for (auto& th : pool) th.join();
  return MC_OK;
}

}  // extern "C"
