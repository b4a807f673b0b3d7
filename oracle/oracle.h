/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.  CPU restatement of the reference's hot paths,
 * used by tests/, the smoke check and bench.py's CPU-baseline leg as the checker.
 * The product (libmcsim.so) never includes, links or calls anything declared here.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>
#include "../include/mcsim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fdlibm.c: V8 Math.* restated */
double or_sin(double x);
double or_cos(double x);
double or_tan(double x);
double or_pow(double x, double y);
double or_hypot(double a, double b);
double or_js_round(double x);
double or_atan(double x);
double or_atan2(double y, double x);
int or_rem_pio2(double x, double *y);
double or_kernel_sin(double x, double y, int iy);
double or_kernel_cos(double x, double y);
double or_kernel_tan(double x, double y, int iy);
/* fn: 0 sin, 1 cos, 2 tan, 3 pow(x, y), 4 hypot(x, y), 5 atan, 6 atan2(x, y) = Math.atan2(x, y) */
void or_math_batch(int fn, const double *x, const double *y, double *out, int64_t n);

/* engine.c: CuttingEngine / chatterAt / timeDomainLimit restated */
typedef struct {
  /* accumulated over a trial, for the FP64 roofline count and the per-step stats */
  int64_t steps;     /* dynamics steps */
  int64_t visits;    /* (tooth, slice) visits: N*nz per step */
  int64_t engaged;   /* visits that reach the force law (h > 0) */
  double maxF;       /* max |(Fx, Fy)| (st.maxF, src/cutting.js:281) */
  double sumF;       /* sum |(Fx, Fy)| */
} or_counts;

int or_chatter_at(const mc_cell *c, double ap, mc_trial *out, or_counts *cnt);
int or_time_domain_limit(const mc_cell *c, const mc_bisect_opts *o, mc_limit *out);
int or_bracket_limit(const mc_cell *c, const mc_bisect_opts *o, double lo_hint,
                     double hi_hint, double tol, mc_limit *out, int32_t *fell_back);
int64_t or_engine_trace(const mc_cell *c, double ap, int64_t nsteps, double *Fx, double *Fy,
                        double *x, double *y);
/* n cells on nthreads host threads (one cell per thread at a time) */
int or_bisect_batch_mt(const mc_cell *cells, int n, const mc_bisect_opts *o, int nthreads,
                       mc_limit *out);
int or_chatter_batch_mt(const mc_cell *cells, const double *ap, int n, int nthreads,
                        mc_trial *out);

/* cengine.c: the CuttingEngine object kept across calls (src/cutting.js:21-326) */
typedef struct or_engine or_engine;
or_engine *or_engine_create(void);
void or_engine_destroy(or_engine *e);
int or_engine_configure(or_engine *e, const mc_cell *c, double fnMax, mc_engine_info *info);
void or_engine_set_wear(or_engine *e, double Kte, double Kre);
void or_engine_set_work(or_engine *e, const mc_wall *w);
int or_engine_set_engagement(or_engine *e, double ux, double uy, double fz, double tipZ, double ap,
                             const double *sliceR, const double *sliceKap, int32_t *nz_out);
void or_engine_set_end_engagement(or_engine *e, double fza, double rOut, const double *rings);
void or_engine_set_maps(or_engine *e, const float *Hb, const float *Eb);
void or_engine_get_state(const or_engine *e, mc_engine_state *s);
void or_engine_set_state(or_engine *e, const mc_engine_state *s);
int or_engine_run(or_engine *e, int64_t nSteps);
void or_engine_stats(or_engine *e, mc_engine_stats *out, int take);
int64_t or_engine_chip_events(const or_engine *e, mc_chip_event *out, int64_t cap);
void or_engine_chatter_index(const or_engine *e, double out[3]);
void or_engine_read_rings(const or_engine *e, float *hist, float *ts, float *ws, float *wsp);
double or_engine_dt(const or_engine *e);
int or_engine_flutes(const or_engine *e);
double or_engine_run_max_f(const or_engine *e);
double or_engine_dz(const or_engine *e);
int or_engine_max_slices(const or_engine *e);


/* sld.c: zero-order lobes (src/sld.js) */
void or_directional_coeffs(double Kr, double phiSt, double phiEx, double out[4]);
/* writes up to max_pts (rpm, alim) pairs, lobe j's points contiguous; lobe_len[j] =
 * points of lobe j (0..lobes-1); returns total points */
int or_stability_lobes(int N, double Ktc, double Krc, double phiSt, double phiEx, int lobes,
                       double rpmMin, double rpmMax, const mc_mode *mx, int nmx,
                       const mc_mode *my, int nmy, double *rpm, double *alim, int max_pts,
                       int *lobe_len);
double or_alim_at(const double *rpm, const double *alim, const int *lobe_len, int nlobes,
                  double S);
/* Simulator.lobes' composite curve (src/sim.js:270-281), one speed node per entry (modes of
 * node i at mx + i * mstride, my + i * mstride); returns the points (counted only when rpm
 * is NULL) */
int or_composite_lobes(int n, const int *N, const double *Ktc, const double *Krc, const double *phiSt,
                       const double *phiEx, const int *lobes, const double *S0, const double *S1,
                       const mc_mode *mx, const int *nmx, const mc_mode *my, const int *nmy, int mstride,
                       double step, double hi, double *rpm, double *alim);

/* stock.c: Stock (src/stock.js) */
typedef struct or_stock or_stock;
or_stock *or_stock_new(double w, double l, double h, double cell);
void or_stock_free(or_stock *s);
void or_stock_dims(const or_stock *s, int32_t *nx, int32_t *ny, double *hx, double *hy);
void or_stock_set_wall_band(or_stock *s, double band);
float *or_stock_h(or_stock *s);
float *or_stock_err(or_stock *s);
float *or_stock_walld(or_stock *s);
int32_t *or_stock_linecut(or_stock *s);
int32_t *or_stock_linewall(or_stock *s);
double or_stock_stamp(or_stock *s, double cx, double cy, double tipZ, double R, int type,
                      double errVal, int line);
void or_stamp_batch(or_stock *s, const mc_stamp *st, int n, double *vol);
/* mark = {ux, uy, f} or NULL; lut = f32 profile of lutn entries or NULL */
double or_stock_stamp_ex(or_stock *s, double cx, double cy, double tipZ, double R, int type,
                         double errVal, int line, const double *mark, const float *lut,
                         int lutn);
void or_stamp_batch_ex(or_stock *s, const mc_stamp *st, const mc_stamp_opt *opt, int n,
                       const float *luts, const int32_t *lut_off, const int32_t *lut_len,
                       double *vol);
float *or_stock_marks(or_stock *s);
float *or_stock_markf(or_stock *s);
float *or_stock_wallerr(or_stock *s);
void or_stock_mark_wall(or_stock *s, double cx, double cy, double R, double nx, double ny,
                        double val);
void or_stock_dirty(const or_stock *s, int32_t *rect);
void or_stock_clear_dirty(or_stock *s);
float or_stock_block_max(or_stock *s, int32_t bi, int32_t bj);
void or_stock_max_in_disk(or_stock *s, double cx, double cy, double r, double above, double *out);
double or_stock_sample(const or_stock *s, double x, double y);
double or_stock_zbottom(const or_stock *s);
/* engage.c: Simulator.fillEngagement / immersion (src/sim.js:301-348) */
double or_fill_engagement(const or_stock *s, double cx, double cy, double ux, double uy, double R,
                          float *Hb);
int or_immersion(const float *Hb, double tipZ, double *st, double *ex);
/* tipLut (src/assembly.js:38-54): returns n; fills out when not NULL */
int or_tip_lut(int shape, double D, double cornerR, const double *er, int en, double edz,
               float *out);

int64_t or_engine_trace_ex(const mc_cell *c, const mc_engage_spec *s, int64_t nsteps, double *Fx,
                           double *Fy, double *x, double *y);

/* toolModel (src/tool.js:16-32) for a uniform carbide beam: k [N/mm], fn [Hz] */
void or_tool_model(double D, double stickout, double kSpindle, double mHolder, double *k,
                   double *fn);

/* simloop.c: the simulator's sub-step loop, rigid workpiece (src/sim.js:380-445) */
double or_fill_end_engagement(const or_stock *s, double cx, double cy, double tipZ, double ux, double uy,
                              double R, int ball, float *Eb, double *rOut_out);
int or_sim_run(or_stock *act, or_stock *ide, or_engine *e, const mc_substep *sub, int n, double R,
               double fluteLen, double rw, double share, int type, const float *lut, int lutn, double cornerR,
               mc_substep_out *out, double dynBudget, double fz0, mc_frame_out *frames, int nframes, int flex,
               double E, double rho, double kTool, mc_wall_out *walls, const double *edge_r, int edge_n,
               double edge_dz, double *budgetLeft);
/* wall.c: wallAt (src/wall.js:11-112); 1 with the wall fields of *w set, 0 for null */
int or_wall_at(const or_stock *s, double px, double py, double pz, double R, const float *Hb, double ux, double uy,
               double zLoad, double E, double rho, double kTool, mc_wall_out *w);
int or_engine_spt(const or_engine *e);

#ifdef __cplusplus
}
#endif
#endif
