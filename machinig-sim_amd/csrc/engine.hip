// engine.hip — a device-resident CuttingEngine (src/cutting.js:21-326) for gfx950.
//
// The engine object a simulator keeps across its sub-steps lives in device memory: modal
// state, delay line (surf), wall mode, per-tooth chip volumes, statistics, the decimated
// history and the tooth / wall sample rings.  A run of n steps is one workgroup of one
// wavefront per engine; mc_engine_run_batch puts many engines in one launch.
//
// The step loop itself is engine_steps (engine_dev.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_host.h"
#include "mc_math.h"
#include "engine_dev.h"

using namespace mce;

namespace {

__global__ __launch_bounds__(64) void engine_run(EngDev* const* engines) {
  __shared__ EngLds L;
  EngDev& E = *engines[blockIdx.x];
  engine_steps(E, E.nsteps, L);
}

// chatterIndex() over the device rings (src/cutting.js:299-318), one lane
__global__ __launch_bounds__(64) void engine_chatter_index(const EngDev* E, double* out) {
  if (threadIdx.x != 0) return;
  eng_chatter_index(*E, out);
}

__global__ void fill_zero_f32(float* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
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

void run_constants(mc_engine* e) {
  eng_run_constants(e->h, e->mat, e->nonlinear != 0, e->wear ? e->wKte : e->mat.Kte, e->wear ? e->wKre : e->mat.Kre);
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
  eng_new_stats(e->h.st);
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
  eng_set_work_null(h);
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
  if (!w) { eng_set_work_null(h); return MC_OK; }
  eng_set_work(h, w->k, w->m, w->zeta, w->nx, w->ny);
  return MC_OK;
}

// setEngagement(u, fz, tipZ, ap), src/cutting.js:98-124
int mc_engine_set_engagement(mc_engine* e, double ux, double uy, double fz, double tipZ, double ap,
                             const double* sliceR, const double* sliceKap, int32_t* nz_out) {
  if (int st = check_engine(e)) return st;
  if (!e->configured) return mch::fail(MC_ERR_ARG, "setEngagement before configure");
  if ((sliceR == nullptr) != (sliceKap == nullptr)) return mch::fail(MC_ERR_ARG, "sliceR and sliceKap go together");
  const int nz = eng_set_engagement(e->h, ux, uy, fz, tipZ, ap, sliceR, sliceKap, e->type == MC_TOOL_BALL);
  e->engaged = true;
  if (nz_out) *nz_out = nz;
  return MC_OK;
}

// setEndEngagement(fza, rOut, zAt), src/cutting.js:130-143
int mc_engine_set_end_engagement(mc_engine* e, double fza, double rOut, const double* rings) {
  if (int st = check_engine(e)) return st;
  const ToolShape flat{0, 0.0, 0.0, nullptr, 0};
  eng_set_end_engagement(e->h, fza, rOut, rings, flat);
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
  if (take) eng_new_stats(e->h.st);
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

// ---- for the sub-step pipeline (simrun.hip), called with the library lock held
EngDev* mce_engine_mirror(mc_engine* e) { return e && e->configured ? &e->h : nullptr; }
EngDev* mce_engine_device(mc_engine* e) { return e ? e->d : nullptr; }

// the engine's run constants set and its record uploaded for a run on the device
int mce_engine_prepare(mc_engine* e, int64_t nsteps) {
  e->h.db = e->h.dz;   // setEngagement's db, which the process damping constant uses
  run_constants(e);
  e->h.nsteps = nsteps;
  e->h.nev = 0;
  return upload(e);
}

// the record read back after such a run; the pipeline has set an engagement
int mce_engine_finish(mc_engine* e) {
  e->ev_fresh = false;
  e->engaged = true;
  return download(e);
}
