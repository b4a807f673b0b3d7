// path.cpp — the path format and the simulator's sub-step schedule on the host, in the
// reference's arithmetic:
//   mc_path_discretize  program moves -> polyline, arcs cut into chords (src/gcode.js:433-466);
//   mc_sim_schedule     polyline -> the sub-steps Simulator.step visits, with an unbounded
//                       budget or frame by frame (src/sim.js:351-381,389-405).
// Both are sequential by nature (each step's position depends on the running arc length) and
// cost a few operations per sub-step: the schedule of a program is ready long before the
// device pipeline that consumes it (mc_sim_run) would finish its first frame.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_host.h"
#include "mc_math.h"

namespace {

constexpr double JS_PI = 3.141592653589793;
constexpr double CHORD_TOL = 0.001;   // largest chord error of an arc [mm] (src/gcode.js:433)
constexpr double RAPID_RATE = 15000;  // mm/min (src/gcode.js:468)
constexpr int GUARD = 4000;           // loop passes per frame at most (src/sim.js:362)

// Math.hypot of three values as V8 computes it: scale by the largest magnitude, Kahan-
// compensated sum of the squares, sqrt, rescale
double hypot3_js(double a, double b, double c) {
  const double v[3] = {std::fabs(a), std::fabs(b), std::fabs(c)};
  if (std::isinf(v[0]) || std::isinf(v[1]) || std::isinf(v[2])) return INFINITY;
  if (std::isnan(v[0]) || std::isnan(v[1]) || std::isnan(v[2])) return NAN;
  double mx = v[0];
  if (v[1] > mx) mx = v[1];
  if (v[2] > mx) mx = v[2];
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  for (int i = 0; i < 3; i++) {
    const double n = v[i] / mx, summand = n * n - comp, pre = sum + summand;
    comp = (pre - sum) - summand;
    sum = pre;
  }
  return std::sqrt(sum) * mx;
}

// min of several values in the order given, the first smallest kept
double min2(double a, double b) { return b < a ? b : a; }
double min3(double a, double b, double c) { return min2(min2(a, b), c); }

struct Sched {
  const mc_path_point* p;
  int npts;
  double step;
  int N;
  double fo;   // feed override
  std::vector<mc_substep> rows;
  std::vector<int32_t> frames;
  double ux = 1, uy = 0;   // this.u

  // one sub-step on segment k (points k-1 -> k) after segS has advanced to s
  void emit(int k, double segLen, double s, double dtm, double nSubEst, int guard) {
    const mc_path_point &A = p[k - 1], &B = p[k];
    const double f = segLen > 0 ? s / segLen : 1.0;
    mc_substep r{};
    r.x = A.x + (B.x - A.x) * f;
    r.y = A.y + (B.y - A.y) * f;
    r.z = A.z + (B.z - A.z) * f;
    const double dxy = mcm::hypot_js(B.x - A.x, B.y - A.y);
    if (dxy > 1e-6) { ux = (B.x - A.x) / dxy; uy = (B.y - A.y) / dxy; }
    double fza = 0, fzEff = 0;
    if (!B.rapid) {   // a rapid never cuts, so its feeds are not used
      const double Sk = B.S, Fk = B.F * fo;
      fza = (B.z < A.z && Sk >= 1) ? (Fk / (Sk * N)) * ((A.z - B.z) / segLen) : 0.0;
      fzEff = dxy > 1e-6 ? (Fk / (Sk * N)) * (dxy / segLen) : 0.0;
    }
    r.ux = ux;
    r.uy = uy;
    r.dtm = dtm;
    r.fzEff = fzEff;
    r.fza = fza;
    r.dxy = dxy;
    r.rapid = B.rapid ? 1 : 0;
    r.line = B.line;
    r.nSubEst = nSubEst;
    r.guard = guard;
    r.frame = 0;
    rows.push_back(r);
  }

  double seg_len(int k) const {
    return hypot3_js(p[k].x - p[k - 1].x, p[k].y - p[k - 1].y, p[k].z - p[k - 1].z);
  }

  // every sub-step with an unbounded budget (src/sim.js:363-377 with budgetT never binding)
  void unbounded() {
    for (int k = 1; k < npts; k++) {
      const double segLen = seg_len(k);
      const double v = (p[k].rapid ? RAPID_RATE : p[k].F) / 60;
      const double dsMax = p[k].rapid ? 1.0 : step;
      double s = 0;
      while (true) {
        const double ds = min2(dsMax, segLen - s);
        if (segLen - s <= 1e-9) break;
        const double dtm = ds / v;
        s += ds;
        emit(k, segLen, s, dtm, 0.0, 0);
        if (s >= segLen - 1e-9) break;
      }
    }
  }

  // successive Simulator.step(dt_real) frames (src/sim.js:351-377,448)
  void framed(double dt, double speed, int maxFrames) {
    int seg = 1;
    double segS = 0;
    bool done = npts < 2;
    int nfr = 0;
    while (!done && nfr < maxFrames) {
      double budget = dt * speed;
      const int c = std::min(seg, npts - 1);
      const double rate = p[c].rapid ? RAPID_RATE : p[c].F * fo;
      const double nSubEst = std::max(1.0, std::ceil(budget * rate / 60 / step));
      int guard = 0;
      const int first = (int)rows.size();
      while (budget > 1e-9 && !done) {
        if (!(guard < GUARD)) break;
        guard++;
        const double segLen = seg_len(seg);
        const double v = (p[seg].rapid ? RAPID_RATE : p[seg].F * fo) / 60;
        const double dsMax = p[seg].rapid ? 1.0 : step;
        const double ds = min3(dsMax, segLen - segS, v * budget);
        if (segLen - segS <= 1e-9) {
          seg++;
          segS = 0;
          if (seg >= npts) { done = true; seg = npts - 1; }
          continue;
        }
        const double dtm = ds / v;
        budget -= dtm;
        segS += ds;
        emit(seg, segLen, segS, dtm, nSubEst, guard);
        if (segS >= segLen - 1e-9) {
          seg++;
          segS = 0;
          if (seg >= npts) { done = true; seg = npts - 1; }
        }
      }
      if ((int)rows.size() > first) {
        rows[first].frame |= MC_FRAME_FIRST;
        rows.back().frame |= MC_FRAME_LAST;
      }
      frames.push_back(first);
      frames.push_back((int32_t)rows.size() - first);
      nfr++;
    }
  }
};

// one point with the attributes of the move ending at it
void put(std::vector<mc_path_point>& out, double x, double y, double z, const mc_move& m, bool rapid, int move) {
  mc_path_point q;
  q.move = move;
  q.pad = 0;
  q.x = x;
  q.y = y;
  q.z = z;
  q.F = m.F;
  q.S = m.S;
  q.rapid = rapid ? 1 : 0;
  q.line = m.line;
  out.push_back(q);
}

}  // namespace

extern "C" {

int mc_path_discretize(const mc_move* moves, int n, double max_seg, mc_path_point* out, int cap, int* n_out) {
  if (n < 0 || (n > 0 && !moves) || !n_out || cap < 0 || (cap > 0 && !out))
    return mch::fail(MC_ERR_ARG, "bad arguments");
  if (!(max_seg > 0)) return mch::fail(MC_ERR_ARG, "need max_seg > 0");
  std::vector<mc_path_point> pts;
  if (n > 0) put(pts, moves[0].from[0], moves[0].from[1], moves[0].from[2], moves[0], true, 0);
  for (int i = 0; i < n; i++) {
    const mc_move& mv = moves[i];
    if (!mv.arc) {
      put(pts, mv.to[0], mv.to[1], mv.to[2], mv, mv.rapid != 0, i);
      continue;
    }
    const int a = mv.axes[0], b = mv.axes[1], nn = mv.axes[2];
    if (a < 0 || a > 2 || b < 0 || b > 2 || nn < 0 || nn > 2 || a == b || a == nn || b == nn)
      return mch::fail(MC_ERR_ARG, "move %d: bad arc axes", i);
    const double ca = mv.ca, cb = mv.cb;
    const double r = mcm::hypot_js(mv.from[a] - ca, mv.from[b] - cb);
    const double a0 = mcm::atan2_js(mv.from[b] - cb, mv.from[a] - ca);
    const double a1 = mcm::atan2_js(mv.to[b] - cb, mv.to[a] - ca);
    double sweep = a1 - a0;
    if (mv.has_sweep) sweep = mv.sweep;
    else if (mv.cw) { if (sweep >= -1e-9) sweep -= 2 * JS_PI; }
    else if (sweep <= 1e-9) sweep += 2 * JS_PI;
    const double dAng = 2 * mcm::acos_js(std::max(-1.0, 1 - CHORD_TOL / std::max(r, CHORD_TOL)));
    const double as = std::fabs(sweep);
    const double segsD = std::max(std::max(as * r < 1e-9 ? 1.0 : 2.0, std::ceil(as * r / max_seg)), std::ceil(as / dAng));
    if (!(segsD >= 1 && segsD < 1e8)) return mch::fail(MC_ERR_ARG, "move %d: arc of %g chords", i, segsD);
    const int segs = (int)segsD;
    for (int k = 1; k <= segs; k++) {
      double p[3];
      if (k == segs) {   // the last chord ends on the move's end exactly
        p[0] = mv.to[0]; p[1] = mv.to[1]; p[2] = mv.to[2];
      } else {
        const double ang = a0 + sweep * k / segs;
        p[a] = ca + r * mcm::cos_js(ang);
        p[b] = cb + r * mcm::sin_js(ang);
        p[nn] = mv.from[nn] + (mv.to[nn] - mv.from[nn]) * k / segs;
      }
      put(pts, p[0], p[1], p[2], mv, mv.rapid != 0, i);
    }
  }
  if (pts.size() > 0x7fffffff) return mch::fail(MC_ERR_ARG, "too many points");
  *n_out = (int)pts.size();
  if (!out && cap == 0) return MC_OK;   // size query
  if ((size_t)cap < pts.size()) return mch::fail(MC_ERR_ARG, "the path needs %zu points", pts.size());
  std::copy(pts.begin(), pts.end(), out);
  return MC_OK;
}

int mc_sim_schedule(const mc_path_point* pts, int npts, const mc_schedule_opts* o, mc_substep* out, int cap,
                    int* n_sub, int32_t* frames, int frame_cap, int* n_frames) {
  if (npts < 0 || (npts > 0 && !pts) || !o || !n_sub || cap < 0 || (cap > 0 && !out) || frame_cap < 0 ||
      (frame_cap > 0 && !frames))
    return mch::fail(MC_ERR_ARG, "bad arguments");
  if (!(o->hx > 0) || o->flutes < 1) return mch::fail(MC_ERR_ARG, "need hx > 0 and flutes >= 1");
  if (!(o->dt_real >= 0) || (o->dt_real > 0 && (!(o->speed > 0) || o->max_frames < 0)))
    return mch::fail(MC_ERR_ARG, "need dt_real >= 0, and speed > 0 with frames");
  if (o->dt_real > 0 && !n_frames) return mch::fail(MC_ERR_ARG, "frames need n_frames");
  Sched s;
  s.p = pts;
  s.npts = npts;
  s.step = std::min(o->hx, 0.4) * 0.6;
  s.N = o->flutes;
  s.fo = o->dt_real > 0 ? o->feed_override : 1.0;
  if (o->dt_real > 0) s.framed(o->dt_real, o->speed, o->max_frames);
  else s.unbounded();
  if (s.rows.size() > 0x7fffffff) return mch::fail(MC_ERR_ARG, "too many sub-steps");
  *n_sub = (int)s.rows.size();
  const int nf = (int)(s.frames.size() / 2);
  if (n_frames) *n_frames = o->dt_real > 0 ? nf : 0;
  if (!out && cap == 0) return MC_OK;   // size query
  if ((size_t)cap < s.rows.size()) return mch::fail(MC_ERR_ARG, "the schedule needs %zu sub-steps", s.rows.size());
  if (o->dt_real > 0 && frame_cap < nf) return mch::fail(MC_ERR_ARG, "the schedule needs %d frames", nf);
  std::copy(s.rows.begin(), s.rows.end(), out);
  if (o->dt_real > 0) std::copy(s.frames.begin(), s.frames.end(), frames);
  return MC_OK;
}

}  // extern "C"
