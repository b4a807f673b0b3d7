/*
 * mcsim.h — C ABI of the MI355X milling-physics library (libmcsim.so).
 *
 * Two hot paths of the browser milling simulator are served here, each behind the
 * reference's own JavaScript interface (the N-API addon in machinig-sim_amd/csrc binds
 * these entry points; the JS modules in machinig-sim_amd/js keep the reference's
 * exported signatures):
 *
 *   Path A  time-domain regenerative-chatter stability search
 *           (src/stabtd.js, src/cutting.js, src/frf.js, src/sld.js of the reference)
 *   Path B  heightmap stock removal (src/stock.js, driven by src/sim.js:351-449)
 *
 * Conventions: lengths in mm, forces in N, modal stiffness k in N/mm (the engine converts
 * to SI like src/cutting.js:55), angles in rad, speeds in rpm.  Every entry point takes
 * plain host pointers, returns an mc_status and never throws; mc_last_error() gives the
 * message of the last failure on the calling thread.  The library allocates its own
 * device memory; nothing allocated by it crosses the ABI except opaque handles.
 * All entry points are synchronous.  Calls that overlap on one device are serialised
 * by an internal lock.
 */
#ifndef MCSIM_H
#define MCSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- limits (checked at the boundary; larger inputs return MC_ERR_LIMIT) ---- */
#define MC_BINS        180   /* angular bins of the engagement map (src/cutting.js:15) */
#define MC_MAX_MODES     4   /* modes per axis: fitModes keeps at most 4 (src/frf.js:81) */
#define MC_MAX_FLUTES    8
#define MC_MAX_SLICES   64   /* ceil(fluteLen/dz) <= 49 by src/cutting.js:63-64 */
#define MC_CHATTER_IDX 0.45  /* verdict threshold of chatterAt (src/stabtd.js:10) */

typedef enum {
  MC_OK = 0,
  MC_ERR_ARG = 1,     /* invalid argument (null pointer, negative count, NaN input) */
  MC_ERR_LIMIT = 2,   /* input exceeds one of the MC_MAX_* limits */
  MC_ERR_NODEV = 3,   /* no usable gfx950 device */
  MC_ERR_HIP = 4,     /* a HIP runtime call failed */
  MC_ERR_NOMEM = 5    /* device allocation failed */
} mc_status;

enum { MC_TOOL_FLAT = 0, MC_TOOL_BALL = 1, MC_TOOL_BULL = 2, MC_TOOL_PROFILE = 3 };   /* BULL, PROFILE: mc_sim_run only */

/* ---- Path A data ---- */

/* tool: {N, R, helixDeg, type, fluteLen, runout, runoutPhase} (src/cutting.js:37-44) */
typedef struct {
  int32_t N;            /* flutes */
  int32_t type;         /* MC_TOOL_FLAT | MC_TOOL_BALL */
  double R;             /* radius [mm] */
  double helixDeg;      /* helix angle [deg] */
  double fluteLen;      /* flute length [mm] */
  double runout;        /* radial runout [mm]; chatterAt forces 0 (src/stabtd.js:16) */
  double runoutPhase;   /* [rad] */
} mc_tool;

/* mode {k [N/mm], fn [Hz], zeta} (src/frf.js:1-3) */
typedef struct { double k, fn, zeta; } mc_mode;

/* cutting coefficients (src/materials.js:24-33); absent mc / Ci are 0 */
typedef struct {
  double Ktc, Krc, Kac;   /* shearing [N/mm^2] */
  double Kte, Kre, Kae;   /* edge [N/mm] */
  double mc;              /* size-effect exponent */
  double Ci;              /* process-damping coefficient [N/m] */
} mc_material;

/* One stability cell: the arguments of chatterAt / timeDomainLimit except the depth
 * (src/stabtd.js:14,24).  modeSet semantics (src/frf.js:6-11) are resolved by the
 * caller: mx/my hold the x and y mode lists (y = x when the reference has no modesY). */
typedef struct {
  mc_tool tool;
  mc_material mat;
  int32_t nmx, nmy;                 /* 1..MC_MAX_MODES */
  mc_mode mx[MC_MAX_MODES];
  mc_mode my[MC_MAX_MODES];
  double S;                         /* spindle speed [rpm] */
  double phiSt, phiEx;              /* engagement arc [rad]: Hb = 0 inside, -1e3 outside */
  double fz;                        /* feed per tooth [mm] */
  int32_t revs;                     /* revolutions per trial (opts.revs) */
  int32_t nonlinear;                /* opts.nonlinear: size effect + process damping */
} mc_cell;

/* result of one chatterAt trial (src/stabtd.js:14-22, src/cutting.js:299-318) */
typedef struct {
  double index;      /* chatter index sOnce/(sAll+0.05) */
  double amp;        /* sAll [um] */
  double once;       /* sOnce [um] */
  int32_t chatter;   /* index > MC_CHATTER_IDX */
  int32_t nz;        /* axial slices engaged */
  int64_t steps;     /* dynamics steps run */
} mc_trial;

/* timeDomainLimit options (src/stabtd.js:24) */
typedef struct {
  double aMin, aMax;   /* depth range [mm] */
  int32_t iters;       /* bisection iterations */
  int32_t reserved;
} mc_bisect_opts;

/* timeDomainLimit result: {alim, bracket:[lo, hi]}; open = stable at aMax, then
 * alim = lo = aMax and hi = +inf (src/stabtd.js:25) */
typedef struct {
  double alim, lo, hi;
  int32_t open;
  int32_t trials;     /* chatterAt trials run for this cell */
  int64_t steps;      /* dynamics steps summed over the trials */
  int64_t visits;     /* (tooth, slice) visits summed over the trials */
  int64_t engaged;    /* visits that reached the force law (chip thickness > 0) */
} mc_limit;

/* ---- Path A entry points ---- */

/* Initialise the library on a HIP device (0 for the first).  Called implicitly with
 * device 0 by the first entry point that needs the GPU. */
int mc_init(int device);
const char* mc_last_error(void);
/* Writes "<device name> <gcnArch> CUs=<n>" into buf. */
int mc_device_info(char* buf, int len);
/* Device time [ms] of the main kernel of the last batch entry point called on this
 * device (the chatter search, the tile pass of mc_stamp_batch, or the sub-step pipeline of
 * mc_sim_run), measured with HIP events on the library's stream around that kernel alone. */
int mc_last_kernel_time(double* ms);
/* Trials of the last search entry point called on this device (mc_bisect_batch,
 * mc_bisect_prior_batch, mc_bisect_bracket_batch): *serial the trials on the serial path (the
 * sum of the results' mc_limit.trials), *executed every trial the kernels ran: the speculated
 * tree nodes off that path included, and the rounds a tree ran before its cell's aMax check
 * came out stable (a number that depends on timing).  No reference counterpart: it measures
 * the speculation a prior saves. */
int mc_last_search_trials(int64_t* executed, int64_t* serial);

/* Launch schedule of the search (library-wide; no reference counterpart).  The trees of
 * the cells whose trials have >= wide_frac x the longest trial's steps (at most wide_max
 * of them) run first as wide tasks on four waves each; waves_per_cu persistent waves per
 * compute unit; use_graph: replay the launch sequence from a captured hipGraph while its
 * shape and buffers repeat.  Defaults {8, 256, 0.5, 1}.  mcsim/autotune.py picks them by
 * timing candidates on the caller's batch. */
typedef struct {
  int32_t waves_per_cu;
  int32_t wide_max;
  double wide_frac;
  int32_t use_graph;
  int32_t reserved[3];
} mc_search_tuning;
int mc_set_search_tuning(const mc_search_tuning* t);
int mc_get_search_tuning(mc_search_tuning* t);

/* chatterAt (src/stabtd.js:14-22) for n (cell, depth) pairs: cells[i] at depth ap[i]. */
int mc_chatter_batch(const mc_cell* cells, const double* ap, int n, mc_trial* out);

/* timeDomainLimit (src/stabtd.js:24-32) for n cells, one persistent launch.
 * Results are identical to the serial reference: same trials, same verdicts. */
int mc_bisect_batch(const mc_cell* cells, int n, const mc_bisect_opts* opts, mc_limit* out);

/* timeDomainLimit guided by a prior: identical results to mc_bisect_batch (same trials on
 * the serial path, same verdicts, same bracket), but the speculative search evaluates the
 * bisection-tree nodes the prior ln(alim) ~ N(mu[i], sig[i]^2) makes most likely to be on
 * the path, so a good prior (the surrogate's prediction and its residual spread) resolves
 * more levels per round.  sig[i] <= 0 means no prior for cell i.  No reference counterpart:
 * it replaces timeDomainLimit (src/stabtd.js:24-32) where a prior is available. */
int mc_bisect_prior_batch(const mc_cell* cells, int n, const mc_bisect_opts* opts, const double* mu,
                          const double* sig, mc_limit* out);

/* Bracketed search: like mc_bisect_batch, but each cell starts from the bracket
 * [lo_hint[i], hi_hint[i]] (e.g. a surrogate's prediction).  Both ends are checked
 * first; if lo_hint chatters or hi_hint is stable the cell falls back to the full
 * timeDomainLimit search.  Bisection stops when hi - lo <= tol_mm (or after iters
 * halvings inside the hint).  fell_back[i] = 1 when the hint failed (may be NULL). */
int mc_bisect_bracket_batch(const mc_cell* cells, int n, const mc_bisect_opts* opts,
                            const double* lo_hint, const double* hi_hint, double tol_mm,
                            mc_limit* out, int32_t* fell_back);

/* Per-step trace of one chatterAt trial: Fx, Fy [N] and x, y [m] after each of the
 * first nsteps steps (CuttingEngine.run, src/cutting.js:168-293).  Any pointer may be
 * NULL.  nsteps may exceed the trial length; the trace stops at the trial's end and
 * *steps_out receives the number of steps written. */
int mc_engine_trace(const mc_cell* cell, double ap, int64_t nsteps,
                    double* Fx, double* Fy, double* x, double* y, int64_t* steps_out);

/* The engine of one simulator sub-step (src/sim.js:252-257,386-416): CuttingEngine
 * configured for the cell's tool (its own flute length, runout and runoutPhase, and fnMax),
 * modes, material and speed, then setEngagement([ux, uy], fz, tipZ, ap) with engine.Hb =
 * Hb (src/cutting.js:38-124), started from rest with the velocity (vx0, vy0) [m/s] set from
 * outside.  Rigid workpiece, no end edges, flat or ball tools. */
typedef struct {
  double ux, uy;       /* unit feed direction */
  double fz;           /* feed per tooth [mm] */
  double tipZ;         /* tool tip height [mm] */
  double ap;           /* engaged axial depth [mm] */
  double vx0, vy0;     /* initial tool velocity [m/s] */
  double fnMax;        /* configure's fnMax [Hz] (0: the modes alone) */
  float Hb[MC_BINS];   /* material height at the cutter periphery per angle bin [mm] */
} mc_engage_spec;

/* Per-step Fx, Fy [N] and x, y [m] of that engine for nsteps steps (CuttingEngine.run,
 * src/cutting.js:146-293).  The cell's arc, fz and revs are not used.  Any output may be
 * NULL; *steps_out receives nsteps. */
int mc_engine_trace_ex(const mc_cell* cell, const mc_engage_spec* eng, int64_t nsteps,
                       double* Fx, double* Fy, double* x, double* y, int64_t* steps_out);

/* ---- A device-resident CuttingEngine (src/cutting.js:21-320) ----
 *
 * The engine object the simulator keeps across sub-steps: its modal state, delay line
 * (surf), wall mode, per-tooth chip volumes, statistics and sample rings live in device
 * memory between calls, so a sequence of configure / setWear / setWork / setEngagement /
 * setEndEngagement / run calls gives the reference's doubles after every run.  Each
 * handle is independent; mc_engine_run_batch steps many engines in one launch. */
#define MC_END_BINS    36    /* angle bins of the end-face map (src/cutting.js:16) */
#define MC_END_RINGS    6    /* radial rings of the end-face map (src/cutting.js:17) */
#define MC_ENGINE_HIST 4096  /* decimated history ring (src/cutting.js:26) */
#define MC_ENGINE_RING 64    /* once-per-tooth and wall sample rings (src/cutting.js:27-31) */

typedef struct mc_engine mc_engine;

/* what configure derives (src/cutting.js:45-80), for the caller's reads */
typedef struct {
  double dt, omega, Ttooth, dz, S;
  int32_t stepsPerTooth, stepsPerRev, decim, piStep, maxSlices, nmx, nmy, pad;
} mc_engine_info;

/* setWork(wall): a 1-DOF workpiece mode along the wall normal (src/cutting.js:88-96) */
typedef struct { double k, m, zeta, nx, ny; } mc_wall;

/* the engine's public scalar fields.  mc_engine_set_state writes x, y, vx, vy, w, wv, t,
 * decim and hist_i (what callers assign: src/stabtd.js:19, src/sim.js:206,427-428,
 * src/validate.js:81-83); the rest are read-only here */
typedef struct {
  double x, y, vx, vy;     /* tool deflection [m] and velocity [m/s] */
  double w, wv;            /* wall deflection [m] and velocity [m/s] */
  double t;                /* engine time [s] */
  double runMeanW;         /* mean wall deflection over the last run [m] */
  double wk, wnx, wny;     /* wall stiffness [N/m] (0: rigid) and normal */
  int64_t stepCount;
  int32_t pr, decim, nz, pad;
  int64_t hist_i, hist_count, ts_i, ts_count, ws_i, ws_count, wsp_i, wsp_count;
} mc_engine_state;

/* the stats object (newStats, src/cutting.js:323-326), in its field order */
typedef struct {
  double n, sumF, maxF, sumFx, sumFy, sumFz, maxFz, sumT, maxT, sumX, sumY;
  double minX, maxX, minY, maxY, sumZc, wZc, sumW, maxW, sumHb, sumB, sumRb, engTeeth;
} mc_engine_stats;

/* a chip event: pushed when a tooth leaves the cut with volume (src/cutting.js:238-241) */
typedef struct { double vol, phi, z, ux, uy, vx, vy; } mc_chip_event;

/* new CuttingEngine() (src/cutting.js:22-35) */
int mc_engine_create(mc_engine** out);
int mc_engine_destroy(mc_engine* e);
/* configure(tool, modal, mat, S, fnMax) (src/cutting.js:39-82): cell->tool (N, type, R,
 * helixDeg, fluteLen, runout [mm], runoutPhase), the mode lists, cell->mat, cell->S and
 * cell->nonlinear; the cell's arc, fz and revs are not used.  info may be NULL. */
int mc_engine_configure(mc_engine* e, const mc_cell* cell, double fnMax, mc_engine_info* info);
/* setWear({Kte, Kre}) (src/cutting.js:85): edge coefficients used instead of the
 * material's from now on */
int mc_engine_set_wear(mc_engine* e, double Kte, double Kre);
/* setWork(wall) (src/cutting.js:88-96); NULL is setWork(null), a rigid workpiece */
int mc_engine_set_work(mc_engine* e, const mc_wall* wall);
/* setEngagement([ux, uy], fz, tipZ, ap) (src/cutting.js:98-124).  sliceR / sliceKap: the
 * radius and edge factor of each slice from the tool's own profile (tool.rAt), at least nz
 * entries each; both NULL: the flat / ball rule.  *nz (may be NULL) receives the slices. */
int mc_engine_set_engagement(mc_engine* e, double ux, double uy, double fz, double tipZ, double ap,
                             const double* sliceR, const double* sliceKap, int32_t* nz);
/* setEndEngagement(fza, rOut, zAt) (src/cutting.js:130-143).  rings: ringR, ringCos, ringZ,
 * ringDs (MC_END_RINGS each, in that order) from the caller's zAt; NULL: zAt = () => 0. */
int mc_engine_set_end_engagement(mc_engine* e, double fza, double rOut, const double* rings);
/* the caller-filled maps engine.Hb (MC_BINS) and engine.Eb (MC_END_BINS x MC_END_RINGS,
 * [bin][ring]); either may be NULL to leave it as it is */
int mc_engine_set_maps(mc_engine* e, const float* Hb, const float* Eb);
int mc_engine_get_state(const mc_engine* e, mc_engine_state* s);
int mc_engine_set_state(mc_engine* e, const mc_engine_state* s);
/* run(nSteps) (src/cutting.js:146-296); setEngagement must have followed configure */
int mc_engine_run(mc_engine* e, int64_t nSteps);
/* n engines, engine i stepping nSteps[i], in one launch; the handles must be distinct */
int mc_engine_run_batch(mc_engine* const* e, const int64_t* nSteps, int n);
/* the stats accumulated since the last take; take != 0 is takeStats() (src/cutting.js:320) */
int mc_engine_read_stats(mc_engine* e, mc_engine_stats* out, int take);
/* chip events pushed by the last run, in order: *n receives their number; at most cap are
 * copied */
int mc_engine_chip_events(const mc_engine* e, mc_chip_event* out, int64_t cap, int64_t* n);
/* chatterIndex() (src/cutting.js:299-318): out[0] index, out[1] amp, out[2] once */
int mc_engine_chatter_index(mc_engine* e, double out[3]);
/* the sample rings: hist (5 x MC_ENGINE_HIST: Fx, Fy, Fz, x, y), toothSamples, wallSamples
 * and wallSamplesPi (2 x MC_ENGINE_RING each: x, y); any may be NULL */
int mc_engine_read_rings(const mc_engine* e, float* hist, float* ts, float* ws, float* wsp);

/* ---- Path B: heightmap stock (src/stock.js) ---- */

typedef struct mc_stock mc_stock;   /* device-resident heightfield + side maps */

/* new Stock(w, l, h, cell) (src/stock.js:9-40): nx = round(w/cell)+1, hx = w/(nx-1),
 * H = zTop = 0, zBottom = -h, wallD = +inf. */
int mc_stock_create(double w, double l, double h, double cell, mc_stock** out);
int mc_stock_destroy(mc_stock* s);
int mc_stock_dims(const mc_stock* s, int32_t* nx, int32_t* ny, double* hx, double* hy);
/* wallBand (src/stock.js:123); 0 by default */
int mc_stock_set_wall_band(mc_stock* s, double wallBand);
/* copy the whole H (f32, row-major k = j*nx + i) between host and device */
int mc_stock_read_h(const mc_stock* s, float* H);
int mc_stock_write_h(mc_stock* s, const float* H);
/* side maps: err [um] (f32), wallD [mm] (f32), lineCut, lineWall (i32) */
int mc_stock_read_side(const mc_stock* s, float* err, float* wallD, int32_t* lineCut,
                       int32_t* lineWall);

/* one Stock.stamp call (src/stock.js:118): flat/ball; mc_stamp_opt adds marks and profiles */
typedef struct {
  double cx, cy, tipZ, R;
  int32_t type;        /* MC_TOOL_FLAT | MC_TOOL_BALL */
  int32_t line;        /* curLine recorded in lineCut / lineWall */
  double errVal;       /* value written to err where material is removed */
} mc_stamp;

/* Apply n stamps in order with the reference's sequential semantics: final H, err,
 * wallD, lineCut and lineWall equal those of n successive Stock.stamp calls.
 * vol (may be NULL) receives each stamp's returned volume [mm^3] (reduction order
 * differs from the reference: agrees to ~1e-12 relative).  Stamps with
 * tipZ >= zTop are skipped (vol 0) like src/stock.js:119. */
int mc_stamp_batch(mc_stock* s, const mc_stamp* stamps, int n, double* vol);

/* The mark and lut arguments of Stock.stamp (src/stock.js:118) for one stamp. */
typedef struct {
  double ux, uy;       /* feed direction (mark.ux, mark.uy) */
  double f;            /* feed per revolution [mm] (mark.f); <= 0: no feed marks (mark = null) */
  int32_t lut;         /* tip profile: index into the batch's tables, -1: none (lut = null) */
  int32_t pad;
} mc_stamp_opt;

/* Stock.stamp with feed marks and tip profiles (src/stock.js:118-164), n stamps in order.
 * opt (may be NULL: no marks, no profiles) holds one mc_stamp_opt per stamp.  Profile t
 * is luts[lut_off[t] .. lut_off[t] + lut_len[t]) (f32 heights above the tip in radial
 * steps of MC_LUT_DR, as tipLut returns); with a profile the cutter surface is
 * tipZ + lut[min(len-1, round(r / MC_LUT_DR))] whatever the type.  Feed marks (markS /
 * markF) are written for flat stamps without a profile.  Everything else as
 * mc_stamp_batch, which this extends. */
int mc_stamp_batch_ex(mc_stock* s, const mc_stamp* stamps, const mc_stamp_opt* opt, int n,
                      const float* luts, const int32_t* lut_off, const int32_t* lut_len,
                      int nluts, double* vol);

/* mc_stamp_batch with the removed volume summed per program line: line_vol[l] = the sum of
 * what Stock.stamp returns (src/stock.js:162) over the batch's stamps with line l, 0 for a
 * line without stamps; every stamp's line must lie in [0, nlines).  This is the simulator's
 * per-segment removed volume (the stamps of one path segment carry its line), summed on the
 * device; its last bits depend on the summation order. */
int mc_stamp_batch_lines(mc_stock* s, const mc_stamp* stamps, int n, double* line_vol, int nlines);

/* mc_stamp_batch without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during a call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The per-stamp records live in a
 * fixed scratch array of the library (MC_NOALLOC_SCRATCH_MB, held by each device the library
 * runs on until the process ends) instead of the stock's grow-only buffers, so at most
 * MC_NOALLOC_MAX_STAMPS stamps per call (more: MC_ERR_LIMIT, nothing applied).  Calls no
 * hipMalloc / hipFree.  Results as mc_stamp_batch.  Every call on a device shares that
 * device's array, which keeps the last call's contents; calls that overlapped on one device
 * would race for it, so the library's call lock serialises them within a process. */
#define MC_NOALLOC_MAX_STAMPS 65536
#define MC_NOALLOC_SCRATCH_MB 13
int mc_stamp_batch_noalloc(mc_stock* s, const mc_stamp* stamps, int n, double* vol);

/* feed-mark maps markS [mm] and markF [mm] (f32; src/stock.js:19-23); either may be NULL */
int mc_stock_read_marks(const mc_stock* s, float* markS, float* markF);

/* Stock.dirty (src/stock.js:81, :156-160): node rectangle {i0, j0, i1, j1} covering the
 * bounding boxes of the stamps that changed a node since the last clear; empty as
 * {INT32_MAX, INT32_MAX, -1, -1} (the reference's {Infinity, Infinity, -1, -1}). */
int mc_stock_dirty(const mc_stock* s, int32_t rect[4]);
int mc_stock_clear_dirty(mc_stock* s);

/* Stock.blockMax over every block (src/stock.js:45-54): highest H of each MC_STOCK_B x
 * MC_STOCK_B node block, bnx = ceil(nx / B) by bny = ceil(ny / B), row-major. */
#define MC_STOCK_B 8
int mc_stock_block_max(const mc_stock* s, float* bmax);

/* Stock.maxInDisk (src/stock.js:57-80) for n queries q[4i..4i+3] = {cx, cy, r, above}
 * (above = -INFINITY for the default): out[3i..3i+2] = {h, x, y}, h = -INFINITY (x = y = 0)
 * when nothing qualifies.  Same value and location as the reference's block walk. */
int mc_stock_max_in_disk(const mc_stock* s, const double* q, int n, double* out);

/* Stock.sample (src/stock.js:85-92): bilinear height at n points, -inf outside. */
int mc_stock_sample(const mc_stock* s, const double* xy, int n, double* out);

/* One engagement query of the simulator's sub-step (src/sim.js:386-388): the cutter of
 * radius R at (cx, cy) fed along the unit vector (ux, uy), tip at tipZ, flutes fluteLen. */
typedef struct {
  double cx, cy, ux, uy, R, tipZ, fluteLen;
} mc_engage_query;

/* Its answers: hmax = fillEngagement's return (src/sim.js:301-319; -INFINITY when no
 * leading bin meets material), ap = min(fluteLen, max(0, hmax - tipZ)) (:388), and
 * immersion(tipZ) (:343-348): imm = 1 with the arc [phiSt, phiEx] [rad], or imm = 0 (null). */
typedef struct {
  double hmax, ap, phiSt, phiEx;
  int32_t imm, pad;
} mc_engage_result;

/* Simulator.fillEngagement + immersion for n queries against the stock's current heights.
 * Hb (may be NULL) receives MC_BINS f32 heights per query (engine.Hb: the stock height on
 * the cutter periphery per angle bin, -inf behind the cutter or off the stock). */
int mc_stock_engagement(const mc_stock* s, const mc_engage_query* q, int n, float* Hb,
                        mc_engage_result* out);

/* ---- Tip profile table, host side ---- */

#define MC_LUT_DR 0.02   /* radial step of the tip-shape table [mm] (src/assembly.js:8) */
enum { MC_TIP_FLAT = 0, MC_TIP_BALL = 1, MC_TIP_BULL = 2, MC_TIP_EDGE = 3 };

/* tipLut(t) (src/assembly.js:38-54): height of the cutter surface above the tip at radius
 * i * MC_LUT_DR, i < n = ceil(D/2 / MC_LUT_DR) + 1.  shape MC_TIP_EDGE uses the imported
 * envelope edge_r[0..edge_n) sampled every edge_dz from the tip (t.edge); cornerR is read
 * for MC_TIP_BULL only.  *n_out receives n; out needs room for cap >= n floats (out may be
 * NULL to ask for n). */
int mc_tip_lut(int32_t shape, double D, double cornerR, const double* edge_r, int32_t edge_n,
               double edge_dz, float* out, int32_t cap, int32_t* n_out);

/* ---- The simulator's sub-step pipeline on the device (src/sim.js:351-449) ----
 *
 * Simulator.step per sub-step: the engagement map from the actual stock (fillEngagement,
 * and fillEndEngagement while descending), the axial depth, the engine set up and run for
 * the sub-step's share of steps (or its vibration halved out of the cut), then the
 * commanded stamp at p and the actual stamp at p plus the deflection, with the feed mark.
 * One wavefront per program runs its whole sub-step list without returning to the host;
 * mc_sim_run puts many programs in one launch.  The workpiece is rigid (workpieceFlex off,
 * so setWork(null); the thin-wall form, flex, is switched off in this build), the flank wear
 * fixed for the call (the engine's setWear and rw), and
 * the tool flat, ball, bull-nose or an imported edge profile (its cutterRadiusAt in the
 * engine's slices, its tip table in the stamps and the end face, as the simulator sets them
 * up, src/sim.js:190-218). */

/* one sub-step, as the simulator's schedule produces it (src/sim.js:353-381,389-405) */
typedef struct {
  double x, y, z;     /* tool position p [mm] */
  double ux, uy;      /* the feed direction the simulator holds (this.u) */
  double dtm;         /* machine time of the sub-step [s] */
  double fzEff;       /* feed per tooth along the move's xy path [mm] (0 on vertical moves) */
  double fza;         /* axial feed per tooth while descending [mm] (0 otherwise) */
  double dxy;         /* xy length of the move [mm]: below 1e-6 a vertical move */
  int32_t rapid;      /* a G0 move */
  int32_t line;       /* program line (Stock.curLine) */
  double nSubEst;     /* its frame's sub-step estimate (src/sim.js:358) */
  int32_t guard;      /* the frame loop's count with this sub-step, segment switches included (:362) */
  int32_t frame;      /* MC_FRAME_FIRST: the frame's dynamics budget starts here (:360);
                         MC_FRAME_LAST: the frame ends after it (collect, :450) */
} mc_substep;
enum { MC_FRAME_FIRST = 1, MC_FRAME_LAST = 2 };

/* what the engine gave a frame (Simulator.collect, src/sim.js:644-663): its takeStats() and,
   after it, its chatterIndex() */
typedef struct {
  mc_engine_stats stats;   /* takeStats() at the frame's end (src/cutting.js:320) */
  double index, amp, once; /* chatterIndex() (src/cutting.js:299-318) */
  int32_t end;             /* sub-steps of the call done when the frame ended */
  int32_t pad;
} mc_frame_out;

/* ---- The path format and the sub-step schedule (host) ----
 *
 * The program's moves become a polyline (arcs cut into chords, src/gcode.js:433-466) and
 * the polyline the sub-steps Simulator.step visits (src/sim.js:351-381,389-405), in the
 * reference's arithmetic (V8's hypot, atan2, acos, cos, sin), so that a JavaScript or C
 * caller hands mc_sim_run the simulator's own sub-steps. */

/* one program move, as the reference's parser emits it (src/gcode.js moves) */
typedef struct {
  double from[3], to[3];   /* end points [mm] */
  double F, S;             /* feed [mm/min] and speed [rpm] */
  int32_t rapid, line;     /* a G0 move; program line */
  int32_t arc;             /* 1: a circular move about (ca, cb) in the plane of axes[0], axes[1] */
  int32_t cw;              /* clockwise (G2) */
  int32_t axes[3];         /* plane axes a, b and the normal n: 0 x, 1 y, 2 z */
  int32_t has_sweep;       /* 1: sweep set (by cutter compensation) instead of derived */
  double ca, cb, sweep;
} mc_move;

/* one polyline point, carrying the attributes of the segment that ends at it */
typedef struct {
  double x, y, z, F, S;
  int32_t rapid, line;
  int32_t move;            /* the move whose segment ends here (the first point: move 0's start) */
  int32_t pad;
} mc_path_point;

/* the path discretizer (src/gcode.js:435-466): the first move's start (as a rapid), then
 * each line move's end, each arc's chords (segs = max(2, or 1 for a null arc;
 * ceil(|sweep| r / max_seg); ceil(|sweep| / dAng)), dAng the angle whose chord lies 0.001 mm
 * inside the arc; the last chord ends on the move's end).  out NULL and cap 0 ask for n. */
int mc_path_discretize(const mc_move* moves, int n, double max_seg, mc_path_point* out, int cap, int* n_out);

typedef struct {
  double hx;             /* stock grid spacing [mm]: stampStep = min(hx, 0.4) * 0.6 (src/sim.js:355) */
  double dt_real;        /* frame length [s] of Simulator.step(dt); 0 for one unbounded run */
  double speed;          /* playback speed (machine seconds per real second) */
  double feed_override;
  int32_t flutes;        /* the tool's flutes (feed per tooth) */
  int32_t max_frames;    /* frames at most (dt_real > 0) */
} mc_schedule_opts;

/* the sub-steps of a polyline (src/sim.js:351-381,389-405).  dt_real 0: every sub-step of
 * the path with an unbounded budget (no frames; nSubEst, guard, frame 0).  dt_real > 0:
 * successive frames of dt_real * speed machine seconds, the frame loop's guard of 4000
 * passes, each frame's sub-step estimate and its first / last flags, and frames[2 f],
 * frames[2 f + 1] = the first sub-step and the sub-step count of frame f.  out NULL and cap
 * 0 (frames NULL, frame_cap 0) ask for the counts. */
int mc_sim_schedule(const mc_path_point* pts, int npts, const mc_schedule_opts* opts, mc_substep* out, int cap,
                    int* n_sub, int32_t* frames, int frame_cap, int* n_frames);

/* what a sub-step did */
typedef struct {
  double dx, dy;      /* deflection added to the actual stamp [mm] */
  double ap;          /* engaged axial depth [mm] */
  double errVal;      /* vibration written by the actual stamp [um] */
  double vol;         /* volume the actual stamp removed [mm^3] */
  double maxF;        /* largest |(Fx, Fy)| of the sub-step's run [N] (0 without a run) */
  double phiSt, phiEx; /* the engagement arc the sub-step shows (src/sim.js:421-423): immersion
                          (:343-348) when it cuts deeper than 0.01 mm and a leading bin stands
                          above tipZ + 0.1, [0, pi] for an end cut alone; 0, 0 otherwise */
  int32_t nsteps;     /* engine steps run (0: not cutting) */
  int32_t cut;        /* 1 when the sub-step cut */
  int32_t arc;        /* 1 when the sub-step sets phiSt / phiEx (the simulator keeps the
                         previous arc otherwise); the report's per-line engagement reads it */
  int32_t pad;
} mc_substep_out;

/* the thin wall the simulator's wall model found at a sub-step (wallAt, src/wall.js:11-112), what
   it passed to setWork, and the wall-error mark it left after the stamp (src/sim.js:443) */
typedef struct {
  double k, m, fn, zeta;   /* setWork's argument: stiffness [N/mm], tip-equivalent mass [kg],
                              natural frequency [Hz] (capped at MC_WALL_FN_CAP), damping ratio */
  double nx, ny;           /* the wall normal (the thinnest crossing near the contact normal) */
  double t, tRoot;         /* thickness at the load height and at the root [mm] */
  double hWall, a;         /* free height above the higher floor, load height above it [mm] */
  double bEff, base;       /* effective plate width [mm], the wall's base height [mm] */
  double cx, cy;           /* the contact point [mm] */
  double I;                /* second moment at the load height [mm^4] */
  double markVal;          /* marked: markWall's value max(0, runMeanW 1e6) [um] */
  double wnx, wny;         /* marked: the engine's wall normal markWall used */
  int32_t found;           /* wallAt gave a wall (the fields above k..I are set); 0: null */
  int32_t marked;          /* the actual stock's wall-error map was marked after the stamp:
                              markWall(x + dx, y + dy, R, wnx, wny, markVal) (src/stock.js:100-113) */
} mc_wall_out;
enum { MC_WALL_FN_CAP = 5000 };   /* Hz (src/wall.js:9): the engine's fnMax with the wall model on */

typedef struct {
  mc_stock* actual;        /* Simulator.stock: cut by the deflected tool */
  mc_stock* ideal;         /* Simulator.ideal: the commanded cut */
  mc_engine* engine;       /* configured for this program's tool and speed, with its setWear */
  const mc_substep* sub;   /* nsub sub-steps, in order */
  mc_substep_out* out;     /* nsub results, or NULL */
  double R;                /* tool radius [mm] */
  double fluteLen;         /* [mm] */
  double rw;               /* radial flank wear taken off the actual stamp [mm] (src/sim.js:439) */
  double share;            /* with dynBudget <= 0: the steps a sub-step may run besides its own
                              time, a fixed share of the frame budget (src/sim.js:410-412) */
  int32_t nsub;
  int32_t type;            /* MC_TOOL_FLAT | MC_TOOL_BALL | MC_TOOL_BULL | MC_TOOL_PROFILE (an
                              imported edge, edge_r below) */
  const float* lut;        /* MC_TOOL_BULL, MC_TOOL_PROFILE: the tool's tipLut (mc_tip_lut), lutn entries */
  int32_t lutn;
  int32_t nframes;         /* capacity of frames */
  double cornerR;          /* MC_TOOL_BULL: corner radius [mm] (cutterRadiusAt, src/assembly.js:23-30) */
  double dynBudget;        /* > 0: the frame's dynamics budget cfg.dynBudget (src/sim.js:50,360): each
                              cutting sub-step runs max(1, min(its own steps, floor(share))) steps
                              with share = max(stepsPerTooth / 4, dynLeft / cost / max(1, nSubEst -
                              guard + 1)), cost = N max(1, nz + end rings), and dynLeft falls by
                              steps x cost (:409-414) */
  double fz;               /* this.fz when the call starts: the feed mark until a sub-step cuts */
  mc_frame_out* frames;    /* NULL, or room for the frames (MC_FRAME_LAST sub-steps) of the call:
                              at each frame's end the engine's stats are taken into it and its
                              chatter index computed */
  int32_t flex;            /* must be 0 in this build: the device form of the thin-wall model is
                              switched off (MC_ERR_LIMIT); the oracle restates it (oracle/wall.c).
                              Meant as: 1: the thin-wall model (cfg.workpieceFlex, src/sim.js:407): each
                              sub-step cutting deeper than 0.01 mm asks wallAt for the wall at the
                              load height tipZ + ap / 2 and passes it to setWork; out of the cut a
                              wall whose deflection has decayed below 1e-7 is dropped (:429).  The
                              engine is then configured with fnMax MC_WALL_FN_CAP (:255).
                              0: a rigid workpiece (setWork(null)) */
  int32_t pad;
  double wallE;            /* flex: the material's Young's modulus [N/mm^2] (mat.E) */
  double wallRho;          /* flex: its density [kg/m^3] (mat.rho) */
  double kTool;            /* flex: the tool's static stiffness [N/mm] (toolStaticK, src/sim.js:250) */
  mc_wall_out* walls;      /* flex: NULL, or nsub records of the wall model per sub-step */
  const double* edge_r;    /* MC_TOOL_PROFILE: the imported cutting-edge envelope, edge_n >= 2 radii
                              [mm] every edge_dz from the tip; the engine's slices take
                              cutterRadiusAt = sampleEdge (src/assembly.js:22-35, src/cutting.js:
                              115-119), the stamps and the end face the tip table lut */
  double edge_dz;
  int32_t edge_n;
  int32_t pad3;
  double* budgetLeft;      /* dynBudget > 0: NULL, or in and out.  In: the frame's budget left when the
                              call starts, taken unless the call's first sub-step starts a frame (a
                              call that begins inside a frame: the simulator re-configures the engine
                              on a spindle-speed change, ensureEngine, src/sim.js:252-257,382, so a
                              caller ends the call before that sub-step, configures, and carries the
                              budget into the next call).  Out: the budget left after the call's last
                              sub-step.  NULL: the full budget when the call starts */
} mc_sim_program;

/* n programs, one workgroup each, one launch; the stocks and engines must be distinct */
int mc_sim_run(const mc_sim_program* progs, int n);

/* ---- V8's Math.* on the host (no reference counterpart as an entry point) ----
 * The fdlibm-method functions the library's kernels use, for host code that must give the
 * reference's doubles (the path discretizer's arcs, src/gcode.js:443-460): out[i] =
 * f(x[i]) or f(x[i], y[i]) (atan2(x, y) is Math.atan2(x[i], y[i])). */
enum { MC_JS_SIN = 0, MC_JS_COS = 1, MC_JS_TAN = 2, MC_JS_ATAN = 3, MC_JS_ATAN2 = 4, MC_JS_ACOS = 5,
       MC_JS_POW = 6, MC_JS_HYPOT = 7 };
int mc_js_math(int32_t fn, const double* x, const double* y, double* out, int64_t n);

/* ---- Analytic (zero-order) stability lobes, host side ---- */

/* stabilityLobes parameters (src/sld.js:23): flutes, Ktc, Krc, the engagement arc, the
 * lobe count, the speed window and the modes (k in N/mm) */
typedef struct {
  int32_t N, lobes;
  double Ktc, Krc, phiSt, phiEx, rpmMin, rpmMax;
  int32_t nmx, nmy;
  mc_mode mx[MC_MAX_MODES];
  mc_mode my[MC_MAX_MODES];
} mc_lobe_args;

/* directionalCoeffs(Kr, phiSt, phiEx) (src/sld.js:6-15): out = [axx, axy, ayx, ayy] */
int mc_directional_coeffs(double Kr, double phiSt, double phiEx, double* out);
/* stabilityLobes (src/sld.js:23-61): the lobes' points, lobe after lobe (lobes with no
 * point are left out), each sorted by speed; lobe_len[j] points in lobe j (0 past the
 * last); *n_out points written (at most max_pts) */
int mc_stability_lobes(const mc_lobe_args* a, double* rpm, double* alim, int max_pts, int* lobe_len, int* n_out);
/* alimAt (src/sld.js:64-76): the lowest lobe covering S, linearly interpolated; +inf if none */
int mc_alim_at(const double* rpm, const double* alim, const int* lobe_len, int nlobes, double S, double* out);
/* Simulator.lobes' composite curve for a speed-dependent tool tip (src/sim.js:270-281, the
 * loop that replaces): node i's diagram, stabilityLobes(nodes[i]) with that node's modes
 * (modalNow at the node's centre), read by alimAt at S = rpmMin, rpmMin + step, ... while
 * S < min(hi, rpmMax); a speed no lobe covers reads 1e3.  The points go out node after node
 * (nodes of several curves may share one call).  *n_out = the points; rpm = alim = NULL with
 * max_pts = 0 only asks for that count. */
int mc_composite_lobes(const mc_lobe_args* nodes, int n, double step, double hi, double* rpm, double* alim,
                       int max_pts, int* n_out);
/* alimAt of each cell at its own speed, on the lobes (of the cell's tool, modes, arc and
 * coefficients) over [0.8 Smin, 1.2 Smax] of the cells sharing them; +inf if none covers */
int mc_analytic_limits(const mc_cell* cells, int n, double* out);

/* ---- Learned bracket surrogate (no counterpart in the reference) ----
 * An MLP 24 -> 256 -> 256 -> 256 -> 1 with GELU (tanh form) that predicts log(alim [mm])
 * of a cell from mc_surrogate_features(); its prediction p gives the bracket
 * [0.85 e^p, 1.15 e^p] for mc_bisect_bracket_batch, which verifies the bracket and
 * otherwise runs the plain search.  Trained on the device: BF16 MFMA products with FP32
 * accumulation, FP32 master weights, Adam; the training step is a replayed hipGraph.
 * Parameters, in this order (FP32): W1[24][256] b1[256] W2[256][256] b2[256]
 * W3[256][256] b3[256] W4[256] b4 (W[i][o]: input i, output o). */
#define MC_SURR_IN      24
#define MC_SURR_HIDDEN 256

typedef struct mc_mlp mc_mlp;

typedef struct {
  int32_t epochs;
  int32_t batch;          /* rows per step: a multiple of 64, <= n */
  float lr, lr_final;     /* Adam step size, cosine-decayed from lr to lr_final */
  float beta1, beta2, eps;
  int32_t fit_scaling;    /* 1: set the input / target scaling from the training rows */
  uint64_t seed;          /* shuffling */
} mc_train_opts;

/* the 24 raw inputs of n cells (row-major n x 24; layout in mlp.hip) */
int mc_surrogate_features(const mc_cell* cells, int n, float* out);
int mc_mlp_create(uint64_t seed, mc_mlp** out);
int mc_mlp_destroy(mc_mlp* m);
int mc_mlp_nparams(void);
/* read (write = 0) or replace (write = 1) the FP32 parameters, in the order above */
int mc_mlp_params(mc_mlp* m, float* params, int write);
/* inputs are used as (x - mean) / std, targets as (t - y_mean) / y_std */
int mc_mlp_set_scaling(mc_mlp* m, const float* x_mean, const float* x_std, float y_mean, float y_std);
int mc_mlp_get_scaling(const mc_mlp* m, float* x_mean, float* x_std, float* y_mean, float* y_std);
/* gradient of loss = sum((y - t_scaled)^2) / (2n) over n rows, in parameter order */
int mc_mlp_grads(mc_mlp* m, const float* X, const float* t, int n, float* grads, float* loss);
/* Adam training on n rows (X n x 24, t = log alim); epoch_loss (may be NULL) gets the mean
 * scaled loss of each epoch.  mc_last_kernel_time() then gives the training time. */
int mc_mlp_train(mc_mlp* m, const float* X, const float* t, int n, const mc_train_opts* o, float* epoch_loss);
/* predicted log(alim) of n rows */
int mc_mlp_predict(mc_mlp* m, const float* X, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif /* MCSIM_H */
