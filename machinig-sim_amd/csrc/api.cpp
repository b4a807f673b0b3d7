// api.cpp — the C ABI of libmcsim.so (include/mcsim.h): argument checks, the per-cell
// constants of CuttingEngine.configure (src/cutting.js:39-82), device buffers, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_dev.h"
#include "mc_host.h"
#include "mc_math.h"

// chatter.hip
int mcd_group_lds_bytes(int nv_cap, int nz_cap);
int mcd_launch_chatter(const mcd::LaunchDev& P, int G, int blocks, hipStream_t stream);
// search.hip
int mcd_search_lds_bytes(int nz_cap);
int mcd_launch_search(const mcd::LaunchDev& P, int blocks, hipStream_t stream);
int mcd_launch_reset_queue(int32_t* q, int n, hipStream_t stream);
int mcd_launch_clear_state(int32_t* s, int n, hipStream_t stream);
int mcd_search_maxspec();
constexpr int SEARCH_MAXP = 2, SEARCH_MAXSLOT = 8;   // slots per wave (a wide task: 4 waves x 2)
// stock.hip
size_t mcs_stampdev_bytes();
size_t mcs_chunkbox_bytes();
int mcs_chunk();
size_t mcs_stampx_bytes();
size_t mcs_cull_bytes();
int mcs_launch_stamps(const mc_stamp* d_in, const mc_stamp_opt* d_opt, int n, const float* luts,
                      const int32_t* lut_off, const int32_t* lut_len, const float* lut_min,
                      void* d_stampdev, void* d_stampx, void* d_cull, void* d_box, int nx, int ny,
                      double hx, double hy, double zTop, double zB, double band, float* H,
                      float* err, float* wallD, int32_t* lineCut, int32_t* lineWall, float* markS,
                      float* markF, int32_t* rect, double* vol, bool vol_lines, hipStream_t stream, hipEvent_t ev0,
                      hipEvent_t ev1);
int mcs_launch_set_rect(int32_t* rect, hipStream_t stream);
int mcs_launch_engagement(const float* H, int nx, int ny, double hx, double hy,
                          const mc_engage_query* d_q, int n, float* d_Hb, mc_engage_result* d_out,
                          hipStream_t stream);
int mcs_launch_block_max(const float* H, int nx, int ny, int bnx, int bny, float* bmax,
                         hipStream_t stream);
int mcs_launch_max_in_disk(const float* H, const float* bmax, int nx, int ny, int bnx, double hx,
                           double hy, const double* d_q, int n, double* d_out, hipStream_t stream);
int mcs_launch_sample(const float* H, int nx, int ny, double hx, double hy, const double* d_xy,
                      int n, double* d_out, hipStream_t stream);
int mcs_launch_fill(float* p, size_t n, float v, hipStream_t stream);
int mcs_scratch_setup(int device);
int mcs_scratch_regions(int device, mc_stamp** in, void** stampdev, void** cull, void** box, double** vol);

namespace {

constexpr double JS_PI = 3.141592653589793;
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return fail(MC_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

double js_max(double a, double b) { return a > b ? a : b; }

// ---- per-device state: stream and grow-only work buffers
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(MC_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    n = bytes;
    return 0;
  }
};

// pinned host staging (grow-only), so that a captured graph can copy from fixed addresses
struct HostBuf {
  void* p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return fail(MC_ERR_NOMEM, "hipHostMalloc(%zu) failed", bytes);
    n = bytes;
    return 0;
  }
};

// the search launch sequence (upload, queue reset, kernel, download) captured as a graph,
// reused while the launch shape and buffers stay the same
struct SearchGraph {
  bool valid = false;
  mcd::LaunchDev P;
  int blocks = 0;
  size_t cell_bytes = 0, task_bytes = 0, out_bytes = 0;
  const void *hc = nullptr, *ht = nullptr, *ho = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
};

struct Ctx {
  int dev = -1;
  int cus = 0;
  std::string name, arch;
  hipStream_t stream = nullptr;   // default priority
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  int64_t last_exec = 0, last_serial = 0;   // trials of the last search call: executed, on the serial path
  DevBuf cells, tasks, out, queue, cstate, ring, hist, trace, geo, gbin;
  HostBuf h_cells, h_tasks, h_out;
  SearchGraph sgraph;
  mc_search_tuning tune = {8, 256, 0.5, 1, {0, 0, 0}};
  std::mutex mu;
};

Ctx g_ctx;
std::mutex g_init_mu;

int ensure_init(int device) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ctx.dev >= 0) {
    if (device >= 0 && device != g_ctx.dev) return fail(MC_ERR_ARG, "already initialised on device %d", g_ctx.dev);
    return MC_OK;
  }
  if (device < 0) device = 0;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device)
    return fail(MC_ERR_NODEV, "no HIP device %d (found %d)", device, count);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(MC_ERR_NODEV, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&g_ctx.ev0));
  HIPCHK(hipEventCreate(&g_ctx.ev1));
  // this device's copy of the no-allocation entry point's scratch array
  if (mcs_scratch_setup(device)) return fail(MC_ERR_HIP, "hipGetSymbolAddress of the stamp scratch failed");
  g_ctx.cus = prop.multiProcessorCount;
  g_ctx.name = prop.name;
  g_ctx.arch = prop.gcnArchName;
  g_ctx.dev = device;
  return MC_OK;
}

int env_int(const char* k, int dflt) {
  const char* v = getenv(k);
  return v && *v ? atoi(v) : dflt;
}

// ---- CuttingEngine.configure for chatterAt's engine (runout forced to 0, src/stabtd.js:16)
int check_cell(const mc_cell& c, int i) {
  if (c.tool.N < 1 || c.tool.N > MC_MAX_FLUTES) return fail(MC_ERR_LIMIT, "cell %d: N=%d outside 1..%d", i, c.tool.N, MC_MAX_FLUTES);
  if (c.nmx < 1 || c.nmx > MC_MAX_MODES || c.nmy < 1 || c.nmy > MC_MAX_MODES)
    return fail(MC_ERR_LIMIT, "cell %d: modes per axis must be 1..%d", i, MC_MAX_MODES);
  if (!(c.tool.R > 0) || !std::isfinite(c.S) || !std::isfinite(c.fz) || c.revs < 0 ||
      !std::isfinite(c.tool.fluteLen) || !std::isfinite(c.tool.helixDeg))
    return fail(MC_ERR_ARG, "cell %d: invalid tool / speed / feed / revs", i);
  for (int j = 0; j < c.nmx; j++)
    if (!(c.mx[j].k > 0) || !(c.mx[j].fn > 0)) return fail(MC_ERR_ARG, "cell %d: mode x%d needs k > 0, fn > 0", i, j);
  for (int j = 0; j < c.nmy; j++)
    if (!(c.my[j].k > 0) || !(c.my[j].fn > 0)) return fail(MC_ERR_ARG, "cell %d: mode y%d needs k > 0, fn > 0", i, j);
  // steps of one trial, spt * N * revs, as make_celldev derives them: the kernels count
  // them in 32-bit integers
  const double S = js_max(100, c.S);
  double fTop = 0;
  for (int j = 0; j < c.nmx; j++) fTop = js_max(fTop, c.mx[j].fn);
  for (int j = 0; j < c.nmy; j++) fTop = js_max(fTop, c.my[j].fn);
  const double spt = js_max(48, std::ceil(60 / (S * c.tool.N) * fTop * 24));
  if (!(spt * c.tool.N * (double)c.revs <= 2147483647.0) || !(spt * c.tool.N <= 16777216.0))
    return fail(MC_ERR_LIMIT, "cell %d: %.0f steps per trial (%.0f per revolution) too many", i,
                spt * c.tool.N * (double)c.revs, spt * c.tool.N);
  return MC_OK;
}

void make_celldev(const mc_cell& c, mcd::CellDev& d, double fnMax = 0) {
  std::memset(&d, 0, sizeof d);
  const int N = c.tool.N;
  const double S = js_max(100, c.S);
  d.omega = S * 2 * JS_PI / 60;
  const double Ttooth = 60 / (S * N);
  double fTop = -INFINITY;
  for (int j = 0; j < c.nmx; j++) fTop = js_max(fTop, c.mx[j].fn);
  for (int j = 0; j < c.nmy; j++) fTop = js_max(fTop, c.my[j].fn);
  d.spt = (int)js_max(48, std::ceil(Ttooth * js_max(fTop, fnMax) * 24));
  d.dt = Ttooth / d.spt;
  d.spr = N * d.spt;
  d.dphi = d.omega * d.dt;
  d.pitch = 2 * JS_PI / N;
  d.binScale = MC_BINS / (2 * JS_PI);
  d.fz = c.fz;
  d.hMax = js_max(0.1, 3 * c.fz);
  d.R = c.tool.R;
  d.tanB = mcm::tan_js(c.tool.helixDeg * JS_PI / 180);
  d.toolFluteLen = c.tool.fluteLen;
  d.ball = c.tool.type == MC_TOOL_BALL;
  d.N = N;
  d.revs = c.revs;
  d.decim = (int)js_max(1, mcm::round_js(d.spt / 40.0));
  d.nmx = c.nmx;
  d.nmy = c.nmy;
  const bool nl = c.nonlinear != 0;
  d.Ktc = c.mat.Ktc; d.Krc = c.mat.Krc; d.Kte = c.mat.Kte; d.Kre = c.mat.Kre;
  d.mc = nl ? c.mat.mc : 0;
  d.Ci = nl ? c.mat.Ci : 0;
  for (int j = 0; j < c.nmx; j++) {
    const double w = 2 * JS_PI * c.mx[j].fn;
    d.kx[j] = c.mx[j].k * 1e3;
    d.mx[j] = d.kx[j] / (w * w);
    d.cx[j] = 2 * c.mx[j].zeta * std::sqrt(d.kx[j] * d.mx[j]);
  }
  for (int j = 0; j < c.nmy; j++) {
    const double w = 2 * JS_PI * c.my[j].fn;
    d.ky[j] = c.my[j].k * 1e3;
    d.my[j] = d.ky[j] / (w * w);
    d.cy[j] = 2 * c.my[j].zeta * std::sqrt(d.ky[j] * d.my[j]);
  }
  for (int i = 0; i < N; i++) d.dR[i] = 0.0 * mcm::cos_js(2 * JS_PI * i / N - 0.0);
  for (int b = 0; b < MC_BINS; b++) {
    const double phi = (b + 0.5) / MC_BINS * 2 * JS_PI;
    d.Hb[b] = (phi >= c.phiSt && phi <= c.phiEx) ? 0.0f : -1e3f;
  }
}

// upper bound of the slice count of any trial depth <= apMax (src/cutting.js:63-64,101)
int nz_bound(const mc_cell& c, double apMax) {
  // dz >= max(0.25, tool.fluteLen/48) for every depth, and maxSlices <= 49
  const double dzMin = js_max(0.25, c.tool.fluteLen / 48);
  const double b = std::ceil(apMax / dzMin);
  if (!(b >= 0)) return 0;
  return (int)std::min(b, 49.0);
}

struct TaskSet {
  std::vector<mcd::CellDev> cells;
  std::vector<mcd::TaskDev> tasks;
  std::vector<double> cost;
  std::vector<int> nzcap;
};

int run_tasks(TaskSet& ts, std::vector<mcd::ResultDev>& res, double* h_trace, int64_t trace_steps) {
  Ctx& X = g_ctx;
  const int n = (int)ts.tasks.size();
  res.assign(n, mcd::ResultDev{});
  if (n == 0) return MC_OK;
  // largest-first order balances uneven cells across the persistent groups
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ts.cost[a] > ts.cost[b]; });
  std::vector<mcd::TaskDev> sorted(n);
  int nz_cap = 1, nv_cap = 1;
  int64_t ring_cap = 1;
  for (int q = 0; q < n; q++) {
    sorted[q] = ts.tasks[order[q]];
    const mcd::CellDev& c = ts.cells[sorted[q].cell];
    const int nzc = std::max(1, ts.nzcap[order[q]]);
    nz_cap = std::max(nz_cap, nzc);
    nv_cap = std::max(nv_cap, nzc * c.N);
    ring_cap = std::max<int64_t>(ring_cap, (int64_t)nzc * c.spr);
  }
  if (nz_cap > MC_MAX_SLICES) return fail(MC_ERR_LIMIT, "a trial needs %d slices (max %d)", nz_cap, MC_MAX_SLICES);
  int G = env_int("MC_GROUP_LANES", 0);
  if (G == 0) G = nv_cap <= 8 * 16 ? 16 : nv_cap <= 8 * 32 ? 32 : 64;
  if (G != 8 && G != 16 && G != 32 && G != 64) return fail(MC_ERR_ARG, "MC_GROUP_LANES must be 8, 16, 32 or 64");
  if (nv_cap > 8 * G) return fail(MC_ERR_LIMIT, "%d visits per step exceed 8 passes of %d lanes", nv_cap, G);
  const int gpw = 64 / G;
  const int wpc = std::max(1, env_int("MC_WAVES_PER_CU", 8));
  int64_t groups = std::min<int64_t>(n, (int64_t)X.cus * wpc * gpw);
  if (trace_steps > 0) groups = 1;
  const int blocks = (int)((groups + gpw - 1) / gpw);
  groups = (int64_t)blocks * gpw;
  const int grp_bytes = mcd_group_lds_bytes(nv_cap, nz_cap);
  if ((size_t)gpw * grp_bytes > 64 * 1024) return fail(MC_ERR_LIMIT, "LDS per workgroup %d B too large", gpw * grp_bytes);

  std::lock_guard<std::mutex> lk(X.mu);
  int st;
  if ((st = X.cells.ensure(ts.cells.size() * sizeof(mcd::CellDev)))) return st;
  if ((st = X.tasks.ensure(n * sizeof(mcd::TaskDev)))) return st;
  if ((st = X.out.ensure(n * sizeof(mcd::ResultDev)))) return st;
  if ((st = X.queue.ensure(256))) return st;
  if ((st = X.ring.ensure((size_t)groups * ring_cap * sizeof(float)))) return st;
  if ((st = X.hist.ensure((size_t)groups * mcd::HIST * sizeof(float2)))) return st;
  if (trace_steps > 0 && (st = X.trace.ensure((size_t)trace_steps * 4 * sizeof(double)))) return st;
  HIPCHK(hipMemcpyAsync(X.cells.p, ts.cells.data(), ts.cells.size() * sizeof(mcd::CellDev), hipMemcpyHostToDevice, X.stream));
  HIPCHK(hipMemcpyAsync(X.tasks.p, sorted.data(), n * sizeof(mcd::TaskDev), hipMemcpyHostToDevice, X.stream));
  HIPCHK(hipMemsetAsync(X.queue.p, 0, sizeof(int32_t), X.stream));
  mcd::LaunchDev P;
  P.cells = (const mcd::CellDev*)X.cells.p;
  P.tasks = (const mcd::TaskDev*)X.tasks.p;
  P.out = (mcd::ResultDev*)X.out.p;
  P.queue = (int32_t*)X.queue.p;
  P.ring = (float*)X.ring.p;
  P.hist = (float2*)X.hist.p;
  P.trace = trace_steps > 0 ? (double*)X.trace.p : nullptr;
  P.trace_steps = trace_steps;
  P.ring_cap = ring_cap;
  P.ntasks = n;
  P.nz_cap = nz_cap;
  P.nv_cap = nv_cap;
  P.grp_bytes = grp_bytes;
  HIPCHK(hipEventRecord(X.ev0, X.stream));
  if (mcd_launch_chatter(P, G, blocks, X.stream)) return fail(MC_ERR_HIP, "chatter kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
  HIPCHK(hipEventRecord(X.ev1, X.stream));
  std::vector<mcd::ResultDev> tmp(n);
  HIPCHK(hipMemcpyAsync(tmp.data(), X.out.p, n * sizeof(mcd::ResultDev), hipMemcpyDeviceToHost, X.stream));
  if (trace_steps > 0) HIPCHK(hipMemcpyAsync(h_trace, X.trace.p, (size_t)trace_steps * 4 * sizeof(double), hipMemcpyDeviceToHost, X.stream));
  HIPCHK(hipStreamSynchronize(X.stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, X.ev0, X.ev1));
  X.last_ms += ms;
  for (int q = 0; q < n; q++) res[order[q]] = tmp[q];
  return MC_OK;
}


// The speculative wave-per-cell search (search.hip): tasks are aMax checks
// (TASK_SINGLE at aMax) and bisection trees (TASK_BISECT / TASK_BRACKET).  The trees of
// the cells with the longest trials (steps >= MC_WIDE_FRAC x the longest, default 0.5; at
// most MC_WIDE_MAX of them, default 256) run as wide tasks, four waves each; the
// persistent workgroups take them first, then one task per wave.
int run_search(TaskSet& ts, std::vector<mcd::ResultDev>& res) {
  Ctx& X = g_ctx;
  const int n = (int)ts.tasks.size();
  res.assign(n, mcd::ResultDev{});
  if (n == 0) return MC_OK;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ts.cost[a] > ts.cost[b]; });
  // wide tasks: the longest trees
  const double wide_frac = X.tune.wide_frac;
  const int wide_max = std::max(0, X.tune.wide_max);
  double smax = 0;
  for (int q = 0; q < n; q++) {
    const mcd::CellDev& c = ts.cells[ts.tasks[q].cell];
    if (ts.tasks[q].mode != mcd::TASK_SINGLE) smax = std::max(smax, (double)c.spt * c.N * c.revs);
  }
  std::vector<int> narrow, wide;
  for (int q = 0; q < n; q++) {
    const int i = order[q];
    const mcd::CellDev& c = ts.cells[ts.tasks[i].cell];
    const bool w = ts.tasks[i].mode != mcd::TASK_SINGLE && (double)c.spt * c.N * c.revs >= wide_frac * smax &&
                   (int)wide.size() < wide_max;
    (w ? wide : narrow).push_back(i);
  }
  const int nn = (int)narrow.size(), nw = (int)wide.size();
  std::vector<mcd::TaskDev> sorted(n);
  int nz_cap = 1, nv_cap = 1;
  int64_t ring_cap = 1, pe_cap = 1;
  for (int q = 0; q < n; q++) {
    const int i = q < nn ? narrow[q] : wide[q - nn];
    sorted[q] = ts.tasks[i];
    const mcd::CellDev& c = ts.cells[sorted[q].cell];
    const int nzc = std::max(1, ts.nzcap[i]);
    nz_cap = std::max(nz_cap, nzc);
    nv_cap = std::max(nv_cap, nzc * c.N);
    ring_cap = std::max<int64_t>(ring_cap, (int64_t)nzc * c.spr);
    pe_cap = std::max<int64_t>(pe_cap, (int64_t)c.spr * c.N);
  }
  const int64_t geo_cap = pe_cap * (nz_cap + mcd_search_maxspec());
  if (geo_cap * 16 >= ((int64_t)1 << 31) || (int64_t)SEARCH_MAXSLOT * ring_cap * 4 >= ((int64_t)1 << 31))
    return fail(MC_ERR_LIMIT, "per-wave delay lines or geometry table exceed 2 GiB");
  const int wpc = std::max(4, X.tune.waves_per_cu);   // 8: 2 per SIMD (the VGPR limit)
  const int64_t want = ((int64_t)nn + 4 * (int64_t)nw + 3) / 4;      // workgroups of four waves
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)X.cus * wpc / 4));
  const int lds = mcd_search_lds_bytes(nz_cap);
  if (lds > 64 * 1024) return fail(MC_ERR_LIMIT, "search LDS %d B too large", lds);
  const int64_t waves = (int64_t)blocks * 4;
  std::lock_guard<std::mutex> lk(X.mu);
  int st;
  if ((st = X.cells.ensure(ts.cells.size() * sizeof(mcd::CellDev)))) return st;
  if ((st = X.tasks.ensure(n * sizeof(mcd::TaskDev)))) return st;
  if ((st = X.out.ensure(n * sizeof(mcd::ResultDev)))) return st;
  if ((st = X.queue.ensure(256))) return st;
  // one verdict slot per cell of ts.cells (tasks index cells by their position there)
  if ((st = X.cstate.ensure(std::max<size_t>(1, ts.cells.size()) * sizeof(int32_t)))) return st;
  if ((st = X.ring.ensure((size_t)waves * SEARCH_MAXSLOT * ring_cap * sizeof(float)))) return st;
  if ((st = X.hist.ensure((size_t)waves * SEARCH_MAXSLOT * mcd::HIST * sizeof(float2)))) return st;
  if ((st = X.geo.ensure((size_t)waves * geo_cap * sizeof(double2)))) return st;
  if ((st = X.gbin.ensure((size_t)waves * geo_cap))) return st;
  const size_t cell_bytes = ts.cells.size() * sizeof(mcd::CellDev), task_bytes = n * sizeof(mcd::TaskDev),
               out_bytes = n * sizeof(mcd::ResultDev);
  if ((st = X.h_cells.ensure(cell_bytes)) || (st = X.h_tasks.ensure(task_bytes)) || (st = X.h_out.ensure(out_bytes))) return st;
  std::memcpy(X.h_cells.p, ts.cells.data(), cell_bytes);
  std::memcpy(X.h_tasks.p, sorted.data(), task_bytes);
  mcd::LaunchDev P;
  std::memset(&P, 0, sizeof P);
  P.geo = (double2*)X.geo.p;
  P.gbin = (unsigned char*)X.gbin.p;
  P.geo_cap = geo_cap;
  P.cells = (const mcd::CellDev*)X.cells.p;
  P.tasks = (const mcd::TaskDev*)X.tasks.p;
  P.out = (mcd::ResultDev*)X.out.p;
  P.queue = (int32_t*)X.queue.p;
  P.cstate = (int32_t*)X.cstate.p;
  if (!P.cstate || X.cstate.n < ts.cells.size() * sizeof(int32_t))
    return fail(MC_ERR_HIP, "search state buffer missing");
  P.ntasks = nn;
  P.wtasks = P.tasks + nn;
  P.wout = P.out + nn;
  P.wqueue = P.queue + 16;
  P.nwide = nw;
  P.ring = (float*)X.ring.p;
  P.hist = (float2*)X.hist.p;
  P.ring_cap = ring_cap;
  P.nz_cap = nz_cap;
  P.nv_cap = nv_cap;
  // the launch sequence: upload, queue reset (a kernel, so that a graph holds no memset
  // node), the search, download.  Launched directly, the two events bracket the search
  // kernel; replayed from a graph, they bracket the graph launch (the copies add ~1 ms)
  auto enqueue = [&](bool events) -> int {
    HIPCHK(hipMemcpyAsync(X.cells.p, X.h_cells.p, cell_bytes, hipMemcpyHostToDevice, X.stream));
    HIPCHK(hipMemcpyAsync(X.tasks.p, X.h_tasks.p, task_bytes, hipMemcpyHostToDevice, X.stream));
    if (mcd_launch_reset_queue(P.queue, 64, X.stream)) return fail(MC_ERR_HIP, "queue reset failed");
    if (mcd_launch_clear_state(P.cstate, (int)ts.cells.size(), X.stream)) return fail(MC_ERR_HIP, "state reset failed");
    if (events) HIPCHK(hipEventRecord(X.ev0, X.stream));
    if (mcd_launch_search(P, blocks, X.stream))
      return fail(MC_ERR_HIP, "search kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    if (events) HIPCHK(hipEventRecord(X.ev1, X.stream));
    HIPCHK(hipMemcpyAsync(X.h_out.p, X.out.p, out_bytes, hipMemcpyDeviceToHost, X.stream));
    return MC_OK;
  };
  SearchGraph& G = X.sgraph;
  bool launched = false;
  if (X.tune.use_graph) {
    const bool same = G.valid && std::memcmp(&G.P, &P, sizeof P) == 0 && G.blocks == blocks &&
                      G.cell_bytes == cell_bytes && G.task_bytes == task_bytes && G.out_bytes == out_bytes &&
                      G.hc == X.h_cells.p && G.ht == X.h_tasks.p && G.ho == X.h_out.p;
    if (!same) {
      G.reset();
      if (hipStreamBeginCapture(X.stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int est = enqueue(false);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(X.stream, &g);
        if (est == MC_OK && ce == hipSuccess && g && hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0) == hipSuccess) {
          G.graph = g;
          G.valid = true;
          G.P = P; G.blocks = blocks; G.cell_bytes = cell_bytes; G.task_bytes = task_bytes; G.out_bytes = out_bytes;
          G.hc = X.h_cells.p; G.ht = X.h_tasks.p; G.ho = X.h_out.p;
        } else {
          if (g) (void)hipGraphDestroy(g);
          G.exec = nullptr;
          (void)hipGetLastError();   // capture not available: launch the sequence directly
        }
      }
    }
    if (G.valid) {
      HIPCHK(hipEventRecord(X.ev0, X.stream));
      HIPCHK(hipGraphLaunch(G.exec, X.stream));
      HIPCHK(hipEventRecord(X.ev1, X.stream));
      launched = true;
    }
  }
  if (!launched && (st = enqueue(true))) return st;
  HIPCHK(hipStreamSynchronize(X.stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, X.ev0, X.ev1));
  X.last_ms += ms;
  const mcd::ResultDev* tmp = static_cast<const mcd::ResultDev*>(X.h_out.p);
  for (int q = 0; q < n; q++) {
    res[q < nn ? narrow[q] : wide[q - nn]] = tmp[q];
    X.last_exec += tmp[q].pad;     // trials the task ran, the speculated ones included
  }
  if (getenv("MC_TASK_TIMES")) {
    FILE* f = fopen(getenv("MC_TASK_TIMES"), "w");
    if (f) {
      for (int q = 0; q < n; q++) {
        const int i = q < nn ? narrow[q] : wide[q - nn];
        fprintf(f, "%d %d %lld %lld %d %d %d\n", ts.tasks[i].cell, ts.tasks[i].mode,
                (long long)tmp[q].t_start, (long long)tmp[q].t_end, tmp[q].trials, tmp[q].pad, q >= nn);
      }
      fclose(f);
    }
  }
  return MC_OK;
}

// Cells the speculative search kernel takes: at most SEARCH_MAXP visits per lane at the
// deepest trial, and a size-effect exponent inside pow_cell's domain.  The others (and all
// cells when MC_SEARCH=serial) run on the one-trial-per-group kernel.
bool spec_cell(const mc_cell& c, const mcd::CellDev& d, double aMax) {
  const char* v = getenv("MC_SEARCH");
  if (v && std::strcmp(v, "serial") == 0) return false;
  const int nzb = nz_bound(c, aMax);
  if (c.tool.N * nzb > SEARCH_MAXP * 64) return false;
  // the kernel addresses a wave's delay lines and geometry table with 32-bit byte offsets
  if ((double)SEARCH_MAXSLOT * std::max(nzb, 1) * d.spr * 4 >= 2147483648.0) return false;
  if ((double)d.spr * d.N * (std::max(nzb, 1) + mcd_search_maxspec()) * 16 >= 2147483648.0) return false;
  const bool use_se = d.mc != 0 && d.mc == d.mc;
  return !use_se || (mcm::pow_cell_ok(d.mc) && d.hMax <= 1e27);
}

void split_cells(const mc_cell* cells, const TaskSet& ts, int n, double aMax, std::vector<int>& spec,
                 std::vector<int>& serial) {
  for (int i = 0; i < n; i++) (spec_cell(cells[i], ts.cells[i], aMax) ? spec : serial).push_back(i);
}

double cell_cost(const mcd::CellDev& c) { return (double)c.spt * c.N * c.revs * c.N; }

void to_limit(const mcd::ResultDev& r, mc_limit& l) {
  l.alim = r.alim; l.lo = r.lo; l.hi = r.hi; l.open = r.open; l.trials = r.trials;
  l.steps = r.steps; l.visits = r.visits; l.engaged = r.engaged;
}

// timeDomainLimit for the cells listed in idx: an aMax-check task and a bisection-tree
// task per cell, combined into the reference's result
int plain_search(const mc_cell* cells, const TaskSet& base, const std::vector<int>& idx,
                 const mc_bisect_opts* o, std::vector<mc_limit>& out, const double* mu = nullptr,
                 const double* sig = nullptr) {
  TaskSet ts;
  ts.cells = base.cells;
  const int m = (int)idx.size();
  for (int q = 0; q < m; q++) {
    const int i = idx[q];
    const mcd::CellDev& c = ts.cells[i];
    const double steps = (double)c.spt * c.N * c.revs;
    mcd::TaskDev a;
    std::memset(&a, 0, sizeof a);
    a.cell = i; a.mode = mcd::TASK_SINGLE; a.ap = o->aMax;
    ts.tasks.push_back(a);
    ts.nzcap.push_back(nz_bound(cells[i], o->aMax));
    // the check is queued ahead of its own tree (and of trees up to 4/3 its size): an open
    // cell's tree then stops as soon as the check has answered
    ts.cost.push_back(steps * 2.0);
    mcd::TaskDev t;
    std::memset(&t, 0, sizeof t);
    t.cell = i; t.mode = mcd::TASK_BISECT; t.iters = o->iters; t.lo = o->aMin; t.hi = o->aMax; t.tol = -1;
    t.aMin = o->aMin; t.aMax = o->aMax;
    if (mu && sig && std::isfinite(mu[i]) && sig[i] > 0 && std::isfinite(sig[i])) { t.mu = mu[i]; t.sig = sig[i]; }
    ts.tasks.push_back(t);
    ts.nzcap.push_back(nz_bound(cells[i], o->aMax));
    ts.cost.push_back(steps * 1.5);
  }
  std::vector<mcd::ResultDev> res;
  int st = run_search(ts, res);
  if (st) return st;
  out.resize(m);
  // cells whose rounds needed more geometry-table columns than the search kernel holds
  // (fell_back == 2) run the whole search on the one-trial-per-group kernel
  TaskSet rs;
  rs.cells = base.cells;
  std::vector<int> redo;
  for (int q = 0; q < m; q++) {
    if (res[2 * q].fell_back != 2 && res[2 * q + 1].fell_back != 2) continue;
    const int i = idx[q];
    mcd::TaskDev t;
    std::memset(&t, 0, sizeof t);
    t.cell = i; t.mode = mcd::TASK_BISECT; t.iters = o->iters; t.aMin = o->aMin; t.aMax = o->aMax;
    rs.tasks.push_back(t);
    rs.nzcap.push_back(nz_bound(cells[i], o->aMax));
    rs.cost.push_back(cell_cost(rs.cells[i]));
    redo.push_back(q);
  }
  std::vector<mcd::ResultDev> rr;
  if (!redo.empty() && (st = run_tasks(rs, rr, nullptr, 0))) return st;
  std::vector<char> done(m, 0);
  for (size_t r = 0; r < redo.size(); r++) {
    to_limit(rr[r], out[redo[r]]);
    done[redo[r]] = 1;
  }
  for (int q = 0; q < m; q++) {
    if (done[q]) continue;
    const mcd::ResultDev& A = res[2 * q];
    const mcd::ResultDev& B = res[2 * q + 1];
    mc_limit& l = out[q];
    std::memset(&l, 0, sizeof l);
    if (!A.chatter) {
      l.alim = o->aMax; l.lo = o->aMax; l.hi = INFINITY; l.open = 1; l.trials = 1;
      l.steps = A.steps; l.visits = A.visits; l.engaged = A.engaged;
    } else {
      l.alim = B.alim; l.lo = B.lo; l.hi = B.hi; l.open = 0; l.trials = 1 + B.trials;
      l.steps = A.steps + B.steps; l.visits = A.visits + B.visits; l.engaged = A.engaged + B.engaged;
    }
  }
  return MC_OK;
}

int prepare(const mc_cell* cells, int n, TaskSet& ts) {
  if (n < 0 || (n > 0 && !cells)) return fail(MC_ERR_ARG, "null cells or negative count");
  ts.cells.resize(n);
  for (int i = 0; i < n; i++) {
    int st = check_cell(cells[i], i);
    if (st) return st;
    make_celldev(cells[i], ts.cells[i]);
  }
  return MC_OK;
}

int check_opts(const mc_bisect_opts* o) {
  if (!o) return fail(MC_ERR_ARG, "null options");
  if (!(o->aMin > 0) || !(o->aMax >= o->aMin) || !std::isfinite(o->aMax) || o->iters < 0 || o->iters > 64)
    return fail(MC_ERR_ARG, "need 0 < aMin <= aMax, 0 <= iters <= 64");
  return MC_OK;
}

// ---- stock
struct StockImpl {
  int nx = 0, ny = 0, bnx = 0, bny = 0;
  double w = 0, l = 0, hx = 0, hy = 0, zTop = 0, zB = 0, band = 0;
  float *H = nullptr, *err = nullptr, *wallD = nullptr, *markS = nullptr, *markF = nullptr;
  int32_t *lineCut = nullptr, *lineWall = nullptr;
  int32_t* rect = nullptr;   // Stock.dirty {i0, j0, i1, j1} on the device
  float* bmax = nullptr;     // block maxima, valid while !bmax_stale
  bool bmax_stale = true;
  DevBuf stamps, opt, sdev, sx, cull, box, vol, xy, smp, lut, lutmeta, q, qout, eq, ehb, eout, strip;
};

}  // namespace

namespace mch {
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
int init() { return ensure_init(-1); }
hipStream_t stream() { return g_ctx.stream; }
std::mutex& lock() { return g_ctx.mu; }
int stock_view(mc_stock* s, StockView* v);
int cus() { return g_ctx.cus; }
void set_last_ms(double ms) { g_ctx.last_ms = ms; }
}  // namespace mch

struct mc_stock : StockImpl {};

int mch::stock_view(mc_stock* s, StockView* v) {
  if (!s || !v) return fail(MC_ERR_ARG, "null stock");
  v->nx = s->nx; v->ny = s->ny; v->hx = s->hx; v->hy = s->hy; v->zTop = s->zTop; v->zBottom = s->zB;
  v->band = s->band;
  v->H = s->H; v->err = s->err; v->wallD = s->wallD; v->markS = s->markS; v->markF = s->markF;
  v->lineCut = s->lineCut; v->lineWall = s->lineWall; v->rect = s->rect;
  s->bmax_stale = true;
  return MC_OK;
}

int mch::stock_strips(mc_stock* s, float** smax, int* snx, int* sny) {
  const int nxs = (s->nx + MCH_STRIP_W - 1) / MCH_STRIP_W, nys = (s->ny + MCH_STRIP_H - 1) / MCH_STRIP_H;
  if (int st = s->strip.ensure((size_t)nxs * nys * 4)) return st;
  *smax = (float*)s->strip.p; *snx = nxs; *sny = nys;
  return MC_OK;
}

extern "C" {

int mc_init(int device) { return ensure_init(device); }

int mc_set_search_tuning(const mc_search_tuning* t) {
  if (!t) return fail(MC_ERR_ARG, "null tuning");
  if (t->waves_per_cu < 4 || t->waves_per_cu > 32 || t->wide_max < 0 || !(t->wide_frac >= 0) || t->wide_frac > 1e9)
    return fail(MC_ERR_ARG, "need 4 <= waves_per_cu <= 32, wide_max >= 0, wide_frac >= 0");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  g_ctx.tune = *t;
  return MC_OK;
}

int mc_get_search_tuning(mc_search_tuning* t) {
  if (!t) return fail(MC_ERR_ARG, "null tuning");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  *t = g_ctx.tune;
  return MC_OK;
}

const char* mc_last_error(void) { return g_err.c_str(); }

int mc_last_kernel_time(double* ms) {
  if (!ms) return fail(MC_ERR_ARG, "null pointer");
  *ms = g_ctx.last_ms;
  return MC_OK;
}

int mc_last_search_trials(int64_t* executed, int64_t* serial) {
  if (!executed || !serial) return fail(MC_ERR_ARG, "null pointer");
  *executed = g_ctx.last_exec;
  *serial = g_ctx.last_serial;
  return MC_OK;
}

int mc_device_info(char* buf, int len) {
  int st = ensure_init(-1);
  if (st) return st;
  if (!buf || len <= 0) return fail(MC_ERR_ARG, "null buffer");
  snprintf(buf, len, "%s %s CUs=%d", g_ctx.name.c_str(), g_ctx.arch.c_str(), g_ctx.cus);
  return MC_OK;
}

int mc_chatter_batch(const mc_cell* cells, const double* ap, int n, mc_trial* out) {
  int st = ensure_init(-1);
  if (st) return st;
  if (n > 0 && (!ap || !out)) return fail(MC_ERR_ARG, "null depth or output array");
  TaskSet ts;
  if ((st = prepare(cells, n, ts))) return st;
  g_ctx.last_ms = 0;
  g_ctx.last_exec = g_ctx.last_serial = 0;
  ts.tasks.resize(n);
  ts.cost.resize(n);
  ts.nzcap.resize(n);
  for (int i = 0; i < n; i++) {
    if (!std::isfinite(ap[i])) return fail(MC_ERR_ARG, "depth %d not finite", i);
    mcd::TaskDev& t = ts.tasks[i];
    std::memset(&t, 0, sizeof t);
    t.cell = i; t.mode = mcd::TASK_SINGLE; t.ap = ap[i];
    ts.nzcap[i] = nz_bound(cells[i], ap[i]);
    ts.cost[i] = cell_cost(ts.cells[i]) * std::max(1, ts.nzcap[i]);
  }
  std::vector<mcd::ResultDev> res;
  if ((st = run_tasks(ts, res, nullptr, 0))) return st;
  for (int i = 0; i < n; i++) {
    out[i].index = res[i].index; out[i].amp = res[i].amp; out[i].once = res[i].once;
    out[i].chatter = res[i].chatter; out[i].nz = res[i].nz; out[i].steps = res[i].steps;
  }
  return MC_OK;
}

}  // extern "C"

namespace {
// the trial counts of a search call (mc_last_search_trials): the results' serial trials;
// the cells in `serial` ran on the one-trial-per-group kernel, which runs exactly those
int finish_counts(const mc_limit* out, int n, const std::vector<int>& serial) {
  int64_t s = 0, x = 0;
  for (int i = 0; i < n; i++) s += out[i].trials;
  for (int i : serial) x += out[i].trials;
  g_ctx.last_serial = s;
  g_ctx.last_exec += x;
  return MC_OK;
}
}  // namespace

extern "C" {

int mc_bisect_batch(const mc_cell* cells, int n, const mc_bisect_opts* opts, mc_limit* out) {
  int st = ensure_init(-1);
  if (st) return st;
  if ((st = check_opts(opts))) return st;
  if (n > 0 && !out) return fail(MC_ERR_ARG, "null output array");
  TaskSet ts;
  if ((st = prepare(cells, n, ts))) return st;
  g_ctx.last_ms = 0;
  g_ctx.last_exec = g_ctx.last_serial = 0;
  std::vector<int> spec, serial;
  split_cells(cells, ts, n, opts->aMax, spec, serial);
  if (!spec.empty()) {
    std::vector<mc_limit> lim;
    if ((st = plain_search(cells, ts, spec, opts, lim))) return st;
    for (size_t q = 0; q < spec.size(); q++) out[spec[q]] = lim[q];
  }
  if (serial.empty()) return finish_counts(out, n, serial);
  TaskSet ss;
  ss.cells = ts.cells;
  for (int i : serial) {
    mcd::TaskDev t;
    std::memset(&t, 0, sizeof t);
    t.cell = i; t.mode = mcd::TASK_BISECT; t.iters = opts->iters; t.aMin = opts->aMin; t.aMax = opts->aMax;
    ss.tasks.push_back(t);
    ss.nzcap.push_back(nz_bound(cells[i], opts->aMax));
    ss.cost.push_back(cell_cost(ts.cells[i]) * (ss.nzcap.back() + opts->iters * 0.25 * ss.nzcap.back()));
  }
  std::vector<mcd::ResultDev> res;
  if ((st = run_tasks(ss, res, nullptr, 0))) return st;
  for (size_t q = 0; q < serial.size(); q++) to_limit(res[q], out[serial[q]]);
  return finish_counts(out, n, serial);
}

int mc_bisect_prior_batch(const mc_cell* cells, int n, const mc_bisect_opts* opts, const double* mu,
                          const double* sig, mc_limit* out) {
  int st = ensure_init(-1);
  if (st) return st;
  if ((st = check_opts(opts))) return st;
  if (n > 0 && (!out || !mu || !sig)) return fail(MC_ERR_ARG, "null prior or output array");
  TaskSet ts;
  if ((st = prepare(cells, n, ts))) return st;
  g_ctx.last_ms = 0;
  g_ctx.last_exec = g_ctx.last_serial = 0;
  std::vector<int> spec, serial;
  split_cells(cells, ts, n, opts->aMax, spec, serial);
  if (!spec.empty()) {
    std::vector<mc_limit> lim;
    if ((st = plain_search(cells, ts, spec, opts, lim, mu, sig))) return st;
    for (size_t q = 0; q < spec.size(); q++) out[spec[q]] = lim[q];
  }
  if (serial.empty()) return finish_counts(out, n, serial);
  // the one-trial-per-group kernel runs the serial search; a prior has nothing to guide
  TaskSet ss;
  ss.cells = ts.cells;
  for (int i : serial) {
    mcd::TaskDev t;
    std::memset(&t, 0, sizeof t);
    t.cell = i; t.mode = mcd::TASK_BISECT; t.iters = opts->iters; t.aMin = opts->aMin; t.aMax = opts->aMax;
    ss.tasks.push_back(t);
    ss.nzcap.push_back(nz_bound(cells[i], opts->aMax));
    ss.cost.push_back(cell_cost(ts.cells[i]) * (ss.nzcap.back() + opts->iters * 0.25 * ss.nzcap.back()));
  }
  std::vector<mcd::ResultDev> res;
  if ((st = run_tasks(ss, res, nullptr, 0))) return st;
  for (size_t q = 0; q < serial.size(); q++) to_limit(res[q], out[serial[q]]);
  return finish_counts(out, n, serial);
}

int mc_bisect_bracket_batch(const mc_cell* cells, int n, const mc_bisect_opts* opts,
                            const double* lo_hint, const double* hi_hint, double tol_mm,
                            mc_limit* out, int32_t* fell_back) {
  int st = ensure_init(-1);
  if (st) return st;
  if ((st = check_opts(opts))) return st;
  if (n > 0 && (!out || !lo_hint || !hi_hint)) return fail(MC_ERR_ARG, "null hint or output array");
  TaskSet ts;
  if ((st = prepare(cells, n, ts))) return st;
  g_ctx.last_ms = 0;
  g_ctx.last_exec = g_ctx.last_serial = 0;
  std::vector<int> spec, serial;
  split_cells(cells, ts, n, opts->aMax, spec, serial);
  if (!spec.empty()) {
    // hinted trees; cells whose hint fails (or is invalid) rerun the plain search
    TaskSet hs;
    hs.cells = ts.cells;
    std::vector<int> hinted, plain;
    for (int i : spec) {
      const bool ok = lo_hint[i] > 0 && hi_hint[i] > lo_hint[i] && hi_hint[i] <= opts->aMax &&
                      lo_hint[i] >= opts->aMin;
      if (!ok) { plain.push_back(i); continue; }
      mcd::TaskDev t;
      std::memset(&t, 0, sizeof t);
      t.cell = i; t.mode = mcd::TASK_BRACKET; t.iters = opts->iters; t.lo = lo_hint[i]; t.hi = hi_hint[i];
      t.tol = tol_mm; t.aMin = opts->aMin; t.aMax = opts->aMax;
      hs.tasks.push_back(t);
      hs.nzcap.push_back(nz_bound(cells[i], hi_hint[i]));
      hs.cost.push_back((double)ts.cells[i].spt * ts.cells[i].N * ts.cells[i].revs);
      hinted.push_back(i);
    }
    std::vector<mcd::ResultDev> hr;
    if ((st = run_search(hs, hr))) return st;
    std::vector<mcd::ResultDev> spent(n);
    for (size_t q = 0; q < hinted.size(); q++) {
      const int i = hinted[q];
      // a round that needed more geometry-table columns than the search kernel holds:
      // the whole bracketed search reruns on the one-trial-per-group kernel below
      if (hr[q].fell_back == 2) { serial.push_back(i); continue; }
      if (hr[q].fell_back) { plain.push_back(i); spent[i] = hr[q]; continue; }
      mc_limit& l = out[i];
      std::memset(&l, 0, sizeof l);
      l.alim = hr[q].alim; l.lo = hr[q].lo; l.hi = hr[q].hi; l.open = 0; l.trials = hr[q].trials;
      l.steps = hr[q].steps; l.visits = hr[q].visits; l.engaged = hr[q].engaged;
      if (fell_back) fell_back[i] = 0;
    }
    std::sort(plain.begin(), plain.end());
    std::vector<mc_limit> pl;
    if (!plain.empty() && (st = plain_search(cells, ts, plain, opts, pl))) return st;
    for (size_t q = 0; q < plain.size(); q++) {
      const int i = plain[q];
      mc_limit l = pl[q];
      l.trials += spent[i].trials; l.steps += spent[i].steps; l.visits += spent[i].visits;
      l.engaged += spent[i].engaged;
      out[i] = l;
      if (fell_back) fell_back[i] = 1;
    }
  }
  if (serial.empty()) return finish_counts(out, n, serial);
  TaskSet ss;
  ss.cells = ts.cells;
  for (int i : serial) {
    mcd::TaskDev t;
    std::memset(&t, 0, sizeof t);
    t.cell = i; t.iters = opts->iters; t.aMin = opts->aMin; t.aMax = opts->aMax;
    t.lo = lo_hint[i]; t.hi = hi_hint[i]; t.tol = tol_mm;
    const bool ok = lo_hint[i] > 0 && hi_hint[i] > lo_hint[i] && hi_hint[i] <= opts->aMax &&
                    lo_hint[i] >= opts->aMin;
    t.mode = ok ? mcd::TASK_BRACKET : mcd::TASK_BISECT;
    ss.tasks.push_back(t);
    ss.nzcap.push_back(nz_bound(cells[i], opts->aMax));
    ss.cost.push_back(cell_cost(ts.cells[i]) * ss.nzcap.back());
  }
  std::vector<mcd::ResultDev> res;
  if ((st = run_tasks(ss, res, nullptr, 0))) return st;
  for (size_t q = 0; q < serial.size(); q++) {
    const int i = serial[q];
    to_limit(res[q], out[i]);
    if (fell_back) fell_back[i] = ss.tasks[q].mode == mcd::TASK_BISECT ? 1 : res[q].fell_back;
  }
  return finish_counts(out, n, serial);
}

int mc_engine_trace(const mc_cell* cell, double ap, int64_t nsteps, double* Fx, double* Fy,
                    double* x, double* y, int64_t* steps_out) {
  int st = ensure_init(-1);
  if (st) return st;
  if (!cell || nsteps < 0 || !std::isfinite(ap)) return fail(MC_ERR_ARG, "bad trace arguments");
  TaskSet ts;
  if ((st = prepare(cell, 1, ts))) return st;
  g_ctx.last_ms = 0;
  g_ctx.last_exec = g_ctx.last_serial = 0;
  const mcd::CellDev& c = ts.cells[0];
  const int64_t total = (int64_t)c.spt * c.N * c.revs;
  const int64_t ns = std::min(nsteps, total);
  ts.tasks.resize(1);
  std::memset(&ts.tasks[0], 0, sizeof(mcd::TaskDev));
  ts.tasks[0].cell = 0; ts.tasks[0].mode = mcd::TASK_SINGLE; ts.tasks[0].ap = ap;
  ts.cost = {1.0};
  ts.nzcap = {nz_bound(*cell, ap)};
  std::vector<double> buf((size_t)std::max<int64_t>(ns, 1) * 4);
  std::vector<mcd::ResultDev> res;
  if ((st = run_tasks(ts, res, buf.data(), std::max<int64_t>(ns, 1)))) return st;
  for (int64_t s = 0; s < ns; s++) {
    if (Fx) Fx[s] = buf[4 * s];
    if (Fy) Fy[s] = buf[4 * s + 1];
    if (x) x[s] = buf[4 * s + 2];
    if (y) y[s] = buf[4 * s + 3];
  }
  if (steps_out) *steps_out = ns;
  return MC_OK;
}

int mc_engine_trace_ex(const mc_cell* cell, const mc_engage_spec* eng, int64_t nsteps, double* Fx,
                       double* Fy, double* x, double* y, int64_t* steps_out) {
  int st = ensure_init(-1);
  if (st) return st;
  if (!cell || !eng || nsteps < 0) return fail(MC_ERR_ARG, "bad trace arguments");
  const mc_engage_spec& e = *eng;
  if (!std::isfinite(e.ux) || !std::isfinite(e.uy) || !std::isfinite(e.fz) || !std::isfinite(e.tipZ) ||
      !std::isfinite(e.ap) || !std::isfinite(e.vx0) || !std::isfinite(e.vy0) || !(e.fnMax >= 0) ||
      !std::isfinite(cell->tool.runout) || !std::isfinite(cell->tool.runoutPhase))
    return fail(MC_ERR_ARG, "bad engagement of the simulator's engine");
  if ((st = check_cell(*cell, 0))) return st;
  TaskSet ts;
  ts.cells.resize(1);
  mcd::CellDev& c = ts.cells[0];
  make_celldev(*cell, c, e.fnMax);
  if (!((double)c.spt * c.N <= 16777216.0)) return fail(MC_ERR_LIMIT, "fnMax gives too many steps per revolution");
  // the simulator's engine: its own feed per tooth and map, the tool's runout
  // (src/cutting.js:42-44, src/sim.js:255,403-404)
  c.fz = e.fz;
  c.hMax = js_max(0.1, 3 * e.fz);
  for (int b = 0; b < MC_BINS; b++) c.Hb[b] = e.Hb[b];
  for (int i = 0; i < c.N; i++)
    c.dR[i] = cell->tool.runout * mcm::cos_js(2 * JS_PI * i / c.N - cell->tool.runoutPhase);
  // one trial long enough for the trace
  const int64_t per_rev = (int64_t)c.spt * c.N;
  const int64_t revs = std::max<int64_t>(1, (nsteps + per_rev - 1) / per_rev);
  if (revs > INT32_MAX || (double)per_rev * revs >= 2147483647.0) return fail(MC_ERR_LIMIT, "trace too long");
  c.revs = (int32_t)revs;
  ts.tasks.resize(1);
  std::memset(&ts.tasks[0], 0, sizeof(mcd::TaskDev));
  mcd::TaskDev& t = ts.tasks[0];
  t.cell = 0; t.mode = mcd::TASK_SINGLE; t.ap = e.ap; t.sim = 1;
  t.ux = e.ux; t.uy = e.uy; t.tipZ = e.tipZ; t.vx0 = e.vx0; t.vy0 = e.vy0;
  ts.cost = {1.0};
  ts.nzcap = {std::max(1, nz_bound(*cell, e.ap))};
  g_ctx.last_ms = 0;
  g_ctx.last_exec = g_ctx.last_serial = 0;
  const int64_t ns = std::max<int64_t>(nsteps, 1);
  std::vector<double> buf((size_t)ns * 4);
  std::vector<mcd::ResultDev> res;
  if ((st = run_tasks(ts, res, buf.data(), ns))) return st;
  for (int64_t s = 0; s < nsteps; s++) {
    if (Fx) Fx[s] = buf[4 * s];
    if (Fy) Fy[s] = buf[4 * s + 1];
    if (x) x[s] = buf[4 * s + 2];
    if (y) y[s] = buf[4 * s + 3];
  }
  if (steps_out) *steps_out = nsteps;
  return MC_OK;
}

// ---------------------------------------------------------------- stock

int mc_stock_create(double w, double l, double h, double cell, mc_stock** out) {
  int st = ensure_init(-1);
  if (st) return st;
  if (!out || !(w > 0) || !(l > 0) || !(h > 0) || !(cell > 0)) return fail(MC_ERR_ARG, "bad stock size");
  const double nxd = mcm::round_js(w / cell) + 1, nyd = mcm::round_js(l / cell) + 1;
  if (nxd * nyd > 1.0e9) return fail(MC_ERR_LIMIT, "stock of %.0f x %.0f nodes too large", nxd, nyd);
  mc_stock* s = new mc_stock();
  s->nx = (int)nxd; s->ny = (int)nyd;
  s->w = w; s->l = l; s->zTop = 0; s->zB = -h;
  s->hx = w / (s->nx - 1); s->hy = l / (s->ny - 1);
  s->bnx = (s->nx + MC_STOCK_B - 1) / MC_STOCK_B; s->bny = (s->ny + MC_STOCK_B - 1) / MC_STOCK_B;
  const size_t n = (size_t)s->nx * s->ny;
  if (hipMalloc(&s->H, n * 4) != hipSuccess || hipMalloc(&s->err, n * 4) != hipSuccess ||
      hipMalloc(&s->wallD, n * 4) != hipSuccess || hipMalloc(&s->lineCut, n * 4) != hipSuccess ||
      hipMalloc(&s->lineWall, n * 4) != hipSuccess || hipMalloc(&s->markS, n * 4) != hipSuccess ||
      hipMalloc(&s->markF, n * 4) != hipSuccess || hipMalloc(&s->rect, 16) != hipSuccess ||
      hipMalloc(&s->bmax, (size_t)s->bnx * s->bny * 4) != hipSuccess) {
    mc_stock_destroy(s);
    return fail(MC_ERR_NOMEM, "stock allocation failed");
  }
  hipStream_t q = g_ctx.stream;
  if (mcs_launch_fill(s->H, n, (float)s->zTop, q) || mcs_launch_fill(s->err, n, 0.0f, q) ||
      mcs_launch_fill(s->wallD, n, INFINITY, q) || mcs_launch_fill((float*)s->lineCut, n, 0.0f, q) ||
      mcs_launch_fill((float*)s->lineWall, n, 0.0f, q) || mcs_launch_fill(s->markS, n, 0.0f, q) ||
      mcs_launch_fill(s->markF, n, 0.0f, q) || mcs_launch_set_rect(s->rect, q) ||
      hipStreamSynchronize(q) != hipSuccess) {
    mc_stock_destroy(s);
    return fail(MC_ERR_HIP, "stock initialisation failed");
  }
  *out = s;
  return MC_OK;
}

int mc_stock_destroy(mc_stock* s) {
  if (!s) return MC_OK;
  for (void* p : {(void*)s->H, (void*)s->err, (void*)s->wallD, (void*)s->lineCut, (void*)s->lineWall,
                  (void*)s->markS, (void*)s->markF, (void*)s->rect, (void*)s->bmax})
    if (p) (void)hipFree(p);
  for (DevBuf* b : {&s->stamps, &s->opt, &s->sdev, &s->sx, &s->cull, &s->box, &s->vol, &s->xy,
                    &s->smp, &s->lut, &s->lutmeta, &s->q, &s->qout, &s->eq, &s->ehb, &s->eout, &s->strip})
    if (b->p) (void)hipFree(b->p);
  delete s;
  return MC_OK;
}

int mc_stock_dims(const mc_stock* s, int32_t* nx, int32_t* ny, double* hx, double* hy) {
  if (!s) return fail(MC_ERR_ARG, "null stock");
  if (nx) *nx = s->nx;
  if (ny) *ny = s->ny;
  if (hx) *hx = s->hx;
  if (hy) *hy = s->hy;
  return MC_OK;
}

int mc_stock_set_wall_band(mc_stock* s, double wallBand) {
  if (!s || !std::isfinite(wallBand)) return fail(MC_ERR_ARG, "bad wallBand");
  s->band = wallBand;
  return MC_OK;
}

int mc_stock_read_h(const mc_stock* s, float* H) {
  if (!s || !H) return fail(MC_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(H, s->H, (size_t)s->nx * s->ny * 4, hipMemcpyDeviceToHost));
  return MC_OK;
}

int mc_stock_write_h(mc_stock* s, const float* H) {
  if (!s || !H) return fail(MC_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(s->H, H, (size_t)s->nx * s->ny * 4, hipMemcpyHostToDevice));
  s->bmax_stale = true;
  return MC_OK;
}

int mc_stock_read_side(const mc_stock* s, float* err, float* wallD, int32_t* lineCut, int32_t* lineWall) {
  if (!s) return fail(MC_ERR_ARG, "null stock");
  const size_t b = (size_t)s->nx * s->ny * 4;
  if (err) HIPCHK(hipMemcpy(err, s->err, b, hipMemcpyDeviceToHost));
  if (wallD) HIPCHK(hipMemcpy(wallD, s->wallD, b, hipMemcpyDeviceToHost));
  if (lineCut) HIPCHK(hipMemcpy(lineCut, s->lineCut, b, hipMemcpyDeviceToHost));
  if (lineWall) HIPCHK(hipMemcpy(lineWall, s->lineWall, b, hipMemcpyDeviceToHost));
  return MC_OK;
}

int mc_stamp_batch(mc_stock* s, const mc_stamp* stamps, int n, double* vol) {
  return mc_stamp_batch_ex(s, stamps, nullptr, n, nullptr, nullptr, nullptr, 0, vol);
}

}  // extern "C"

namespace {
// the stamp batch: vol receives the removed volume per stamp (nlines 0) or per program line
// (nlines > 0 entries, every stamp's line in [0, nlines))
int stamp_batch(mc_stock* s, const mc_stamp* stamps, const mc_stamp_opt* opt, int n, const float* luts,
                const int32_t* lut_off, const int32_t* lut_len, int nluts, double* vol, int nlines) {
  if (!s || n < 0 || (n > 0 && !stamps) || nluts < 0 || (nluts > 0 && (!luts || !lut_off || !lut_len)) || nlines < 0)
    return fail(MC_ERR_ARG, "bad stamp batch");
  if (nlines > 0) {
    if (!vol) return fail(MC_ERR_ARG, "per-line volumes need an output");
    for (int i = 0; i < n; i++)
      if (stamps[i].line < 0 || stamps[i].line >= nlines)
        return fail(MC_ERR_ARG, "stamp %d: line %d outside [0, %d)", i, stamps[i].line, nlines);
  }
  // profile tables: offsets and lengths in range, finite heights, and their minima (the
  // lowest reach of a profiled stamp)
  std::vector<float> lmin((size_t)nluts);
  int64_t lut_total = 0;
  for (int t = 0; t < nluts; t++) {
    if (lut_off[t] < 0 || lut_len[t] < 1 || (int64_t)lut_off[t] + lut_len[t] > (int64_t)INT32_MAX)
      return fail(MC_ERR_ARG, "profile %d: bad offset or length", t);
    lut_total = std::max(lut_total, (int64_t)lut_off[t] + lut_len[t]);
    float m = INFINITY;
    for (int i = 0; i < lut_len[t]; i++) {
      const float v = luts[lut_off[t] + i];
      if (!std::isfinite(v)) return fail(MC_ERR_ARG, "profile %d: entry %d not finite", t, i);
      m = std::min(m, v);
    }
    lmin[t] = m;
  }
  for (int i = 0; i < n; i++) {
    const mc_stamp& p = stamps[i];
    if (!std::isfinite(p.cx) || !std::isfinite(p.cy) || !std::isfinite(p.tipZ) || !(p.R > 0) ||
        (p.type != MC_TOOL_FLAT && p.type != MC_TOOL_BALL))
      return fail(MC_ERR_ARG, "stamp %d: bad position, radius or type", i);
    if (opt && (!std::isfinite(opt[i].ux) || !std::isfinite(opt[i].uy) || std::isnan(opt[i].f) ||
                opt[i].lut < -1 || opt[i].lut >= nluts))
      return fail(MC_ERR_ARG, "stamp %d: bad feed mark or profile index", i);
  }
  if (n == 0) return MC_OK;
  int st;
  const int chunks = (n + mcs_chunk() - 1) / mcs_chunk();
  if ((st = s->stamps.ensure((size_t)n * sizeof(mc_stamp)))) return st;
  if ((st = s->sdev.ensure((size_t)n * mcs_stampdev_bytes()))) return st;
  if ((st = s->cull.ensure((size_t)n * mcs_cull_bytes()))) return st;
  if ((st = s->box.ensure((size_t)chunks * mcs_chunkbox_bytes()))) return st;
  const size_t nvol = nlines > 0 ? (size_t)nlines : (size_t)n;
  if (vol && (st = s->vol.ensure(nvol * sizeof(double)))) return st;
  if (opt) {
    if ((st = s->opt.ensure((size_t)n * sizeof(mc_stamp_opt)))) return st;
    if ((st = s->sx.ensure((size_t)n * mcs_stampx_bytes()))) return st;
    if ((st = s->lut.ensure((size_t)std::max<int64_t>(lut_total, 1) * 4))) return st;
    if ((st = s->lutmeta.ensure((size_t)std::max(nluts, 1) * 12))) return st;
  }
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  hipStream_t q = g_ctx.stream;
  HIPCHK(hipMemcpyAsync(s->stamps.p, stamps, (size_t)n * sizeof(mc_stamp), hipMemcpyHostToDevice, q));
  int32_t *d_off = nullptr, *d_len = nullptr;
  float* d_min = nullptr;
  if (opt) {
    HIPCHK(hipMemcpyAsync(s->opt.p, opt, (size_t)n * sizeof(mc_stamp_opt), hipMemcpyHostToDevice, q));
    d_off = (int32_t*)s->lutmeta.p;
    d_len = d_off + std::max(nluts, 1);
    d_min = (float*)(d_len + std::max(nluts, 1));
    if (nluts > 0) {
      HIPCHK(hipMemcpyAsync(s->lut.p, luts, (size_t)lut_total * 4, hipMemcpyHostToDevice, q));
      HIPCHK(hipMemcpyAsync(d_off, lut_off, (size_t)nluts * 4, hipMemcpyHostToDevice, q));
      HIPCHK(hipMemcpyAsync(d_len, lut_len, (size_t)nluts * 4, hipMemcpyHostToDevice, q));
      HIPCHK(hipMemcpyAsync(d_min, lmin.data(), (size_t)nluts * 4, hipMemcpyHostToDevice, q));
    }
  }
  if (vol && mcs_launch_fill((float*)s->vol.p, nvol * 2, 0.0f, q)) return fail(MC_ERR_HIP, "fill failed");
  if (mcs_launch_stamps((const mc_stamp*)s->stamps.p, opt ? (const mc_stamp_opt*)s->opt.p : nullptr, n,
                        opt ? (const float*)s->lut.p : nullptr, d_off, d_len, d_min, s->sdev.p,
                        opt ? s->sx.p : nullptr, s->cull.p, s->box.p, s->nx, s->ny, s->hx,
                        s->hy, s->zTop, s->zB, s->band, s->H, s->err, s->wallD, s->lineCut, s->lineWall,
                        s->markS, s->markF, s->rect, vol ? (double*)s->vol.p : nullptr, nlines > 0, q, g_ctx.ev0,
                        g_ctx.ev1))
    return fail(MC_ERR_HIP, "stamp kernels failed: %s", hipGetErrorString(hipGetLastError()));
  s->bmax_stale = true;
  if (vol) HIPCHK(hipMemcpyAsync(vol, s->vol.p, nvol * sizeof(double), hipMemcpyDeviceToHost, q));
  HIPCHK(hipStreamSynchronize(q));
  {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1));
    g_ctx.last_ms = ms;
  }
  if (vol) {
    const double area = s->hx * s->hy;
    for (size_t i = 0; i < nvol; i++) vol[i] *= area;
  }
  return MC_OK;
}
}  // namespace

extern "C" {

int mc_stamp_batch_ex(mc_stock* s, const mc_stamp* stamps, const mc_stamp_opt* opt, int n,
                      const float* luts, const int32_t* lut_off, const int32_t* lut_len, int nluts,
                      double* vol) {
  return stamp_batch(s, stamps, opt, n, luts, lut_off, lut_len, nluts, vol, 0);
}

int mc_stamp_batch_lines(mc_stock* s, const mc_stamp* stamps, int n, double* line_vol, int nlines) {
  if (nlines < 1) return fail(MC_ERR_ARG, "need nlines >= 1");
  return stamp_batch(s, stamps, nullptr, n, nullptr, nullptr, nullptr, 0, line_vol, nlines);
}

int mc_stamp_batch_noalloc(mc_stock* s, const mc_stamp* stamps, int n, double* vol) {
  if (!s || n < 0 || (n > 0 && !stamps)) return fail(MC_ERR_ARG, "bad stamp batch");
  if (n > MC_NOALLOC_MAX_STAMPS)
    return fail(MC_ERR_LIMIT, "%d stamps: the no-allocation batch takes at most MC_NOALLOC_MAX_STAMPS = %d", n,
                MC_NOALLOC_MAX_STAMPS);
  for (int i = 0; i < n; i++) {
    const mc_stamp& p = stamps[i];
    if (!std::isfinite(p.cx) || !std::isfinite(p.cy) || !std::isfinite(p.tipZ) || !(p.R > 0) ||
        (p.type != MC_TOOL_FLAT && p.type != MC_TOOL_BALL))
      return fail(MC_ERR_ARG, "stamp %d: bad position, radius or type", i);
  }
  if (n == 0) return MC_OK;
  mc_stamp* d_in;
  void *d_dev, *d_cull, *d_box;
  double* d_vol;
  if (mcs_scratch_regions(g_ctx.dev, &d_in, &d_dev, &d_cull, &d_box, &d_vol))
    return fail(MC_ERR_NODEV, "the library is not initialised on a device");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  hipStream_t q = g_ctx.stream;
  HIPCHK(hipMemcpyAsync(d_in, stamps, (size_t)n * sizeof(mc_stamp), hipMemcpyHostToDevice, q));
  // the volumes start at zero (a fill kernel: no memset node if a caller captures this);
  // stamp_prep rewrites the records and boxes of the n stamps
  if (vol && mcs_launch_fill((float*)d_vol, (size_t)n * 2, 0.0f, q)) return fail(MC_ERR_HIP, "fill failed");
  if (mcs_launch_stamps(d_in, nullptr, n, nullptr, nullptr, nullptr, nullptr, d_dev, nullptr, d_cull, d_box,
                        s->nx, s->ny, s->hx, s->hy, s->zTop, s->zB, s->band, s->H, s->err, s->wallD, s->lineCut,
                        s->lineWall, s->markS, s->markF, s->rect, vol ? d_vol : nullptr, false, q, g_ctx.ev0,
                        g_ctx.ev1))
    return fail(MC_ERR_HIP, "stamp kernels failed: %s", hipGetErrorString(hipGetLastError()));
  s->bmax_stale = true;
  if (vol) HIPCHK(hipMemcpyAsync(vol, d_vol, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, q));
  HIPCHK(hipStreamSynchronize(q));
  {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1));
    g_ctx.last_ms = ms;
  }
  if (vol) {
    const double area = s->hx * s->hy;
    for (int i = 0; i < n; i++) vol[i] *= area;
  }
  return MC_OK;
}

int mc_stock_read_marks(const mc_stock* s, float* markS, float* markF) {
  if (!s) return fail(MC_ERR_ARG, "null stock");
  const size_t b = (size_t)s->nx * s->ny * 4;
  if (markS) HIPCHK(hipMemcpy(markS, s->markS, b, hipMemcpyDeviceToHost));
  if (markF) HIPCHK(hipMemcpy(markF, s->markF, b, hipMemcpyDeviceToHost));
  return MC_OK;
}

int mc_stock_dirty(const mc_stock* s, int32_t rect[4]) {
  if (!s || !rect) return fail(MC_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(rect, s->rect, 16, hipMemcpyDeviceToHost));
  return MC_OK;
}

int mc_stock_clear_dirty(mc_stock* s) {
  if (!s) return fail(MC_ERR_ARG, "null stock");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (mcs_launch_set_rect(s->rect, g_ctx.stream)) return fail(MC_ERR_HIP, "clear kernel failed");
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return MC_OK;
}

// block maxima are recomputed from H after any change: the reference's lazily refreshed
// maxima equal the current block maxima whenever a query reads them
static int refresh_bmax(mc_stock* s, hipStream_t q) {
  if (!s->bmax_stale) return MC_OK;
  if (mcs_launch_block_max(s->H, s->nx, s->ny, s->bnx, s->bny, s->bmax, q))
    return fail(MC_ERR_HIP, "block max kernel failed");
  s->bmax_stale = false;
  return MC_OK;
}

int mc_stock_block_max(const mc_stock* s, float* bmax) {
  if (!s || !bmax) return fail(MC_ERR_ARG, "null argument");
  mc_stock* m = const_cast<mc_stock*>(s);
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  int st;
  if ((st = refresh_bmax(m, g_ctx.stream))) return st;
  HIPCHK(hipMemcpyAsync(bmax, s->bmax, (size_t)s->bnx * s->bny * 4, hipMemcpyDeviceToHost, g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return MC_OK;
}

int mc_stock_max_in_disk(const mc_stock* s, const double* qv, int n, double* out) {
  if (!s || n < 0 || (n > 0 && (!qv || !out))) return fail(MC_ERR_ARG, "bad query arguments");
  for (int i = 0; i < n; i++)
    if (!std::isfinite(qv[4 * i]) || !std::isfinite(qv[4 * i + 1]) || !std::isfinite(qv[4 * i + 2]) ||
        std::isnan(qv[4 * i + 3]))
      return fail(MC_ERR_ARG, "query %d: bad centre, radius or floor", i);
  if (n == 0) return MC_OK;
  mc_stock* m = const_cast<mc_stock*>(s);
  int st;
  if ((st = m->q.ensure((size_t)n * 32))) return st;
  if ((st = m->qout.ensure((size_t)n * 24))) return st;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  hipStream_t q = g_ctx.stream;
  if ((st = refresh_bmax(m, q))) return st;
  HIPCHK(hipMemcpyAsync(m->q.p, qv, (size_t)n * 32, hipMemcpyHostToDevice, q));
  if (mcs_launch_max_in_disk(s->H, s->bmax, s->nx, s->ny, s->bnx, s->hx, s->hy, (const double*)m->q.p, n,
                             (double*)m->qout.p, q))
    return fail(MC_ERR_HIP, "maxInDisk kernel failed");
  HIPCHK(hipMemcpyAsync(out, m->qout.p, (size_t)n * 24, hipMemcpyDeviceToHost, q));
  HIPCHK(hipStreamSynchronize(q));
  return MC_OK;
}

int mc_stock_engagement(const mc_stock* s, const mc_engage_query* qv, int n, float* Hb,
                        mc_engage_result* out) {
  if (!s || n < 0 || (n > 0 && (!qv || !out))) return fail(MC_ERR_ARG, "bad engagement arguments");
  for (int i = 0; i < n; i++) {
    const mc_engage_query& q = qv[i];
    if (!std::isfinite(q.cx) || !std::isfinite(q.cy) || !std::isfinite(q.ux) || !std::isfinite(q.uy) ||
        !std::isfinite(q.R) || std::isnan(q.tipZ) || std::isnan(q.fluteLen))
      return fail(MC_ERR_ARG, "engagement query %d: bad position, direction or radius", i);
  }
  if (n == 0) return MC_OK;
  mc_stock* m = const_cast<mc_stock*>(s);
  int st;
  if ((st = m->eq.ensure((size_t)n * sizeof(mc_engage_query)))) return st;
  if ((st = m->eout.ensure((size_t)n * sizeof(mc_engage_result)))) return st;
  if (Hb && (st = m->ehb.ensure((size_t)n * MC_BINS * 4))) return st;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  hipStream_t q = g_ctx.stream;
  HIPCHK(hipMemcpyAsync(m->eq.p, qv, (size_t)n * sizeof(mc_engage_query), hipMemcpyHostToDevice, q));
  if (mcs_launch_engagement(s->H, s->nx, s->ny, s->hx, s->hy, (const mc_engage_query*)m->eq.p, n,
                            Hb ? (float*)m->ehb.p : nullptr, (mc_engage_result*)m->eout.p, q))
    return fail(MC_ERR_HIP, "engagement kernel failed");
  HIPCHK(hipMemcpyAsync(out, m->eout.p, (size_t)n * sizeof(mc_engage_result), hipMemcpyDeviceToHost, q));
  if (Hb) HIPCHK(hipMemcpyAsync(Hb, m->ehb.p, (size_t)n * MC_BINS * 4, hipMemcpyDeviceToHost, q));
  HIPCHK(hipStreamSynchronize(q));
  return MC_OK;
}

int mc_stock_sample(const mc_stock* s, const double* xy, int n, double* out) {
  if (!s || n < 0 || (n > 0 && (!xy || !out))) return fail(MC_ERR_ARG, "bad sample arguments");
  if (n == 0) return MC_OK;
  mc_stock* m = const_cast<mc_stock*>(s);
  int st;
  if ((st = m->xy.ensure((size_t)n * 16))) return st;
  if ((st = m->smp.ensure((size_t)n * 8))) return st;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  hipStream_t q = g_ctx.stream;
  HIPCHK(hipMemcpyAsync(m->xy.p, xy, (size_t)n * 16, hipMemcpyHostToDevice, q));
  if (mcs_launch_sample(s->H, s->nx, s->ny, s->hx, s->hy, (const double*)m->xy.p, n, (double*)m->smp.p, q))
    return fail(MC_ERR_HIP, "sample kernel failed");
  HIPCHK(hipMemcpyAsync(out, m->smp.p, (size_t)n * 8, hipMemcpyDeviceToHost, q));
  HIPCHK(hipStreamSynchronize(q));
  return MC_OK;
}

}  // extern "C"
