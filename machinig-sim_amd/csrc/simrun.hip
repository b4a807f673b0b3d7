// simrun.hip — the simulator's sub-step loop (src/sim.js:351-449) on the device.
//
// One workgroup per program runs its whole list of sub-steps: per sub-step it samples the
// actual stock on the cutter periphery (fillEngagement, src/sim.js:301-319) and, while
// descending, under the end face (fillEndEngagement, :324-341) into the engine's maps,
// takes the axial depth, sets the engine up and runs the sub-step's share of steps
// (engine_steps, the reference's run) or halves the vibration out of the cut, and stamps
// the commanded stock at p and the actual stock at p plus the deflection (Stock.stamp,
// src/stock.js:118-164), before the next sub-step reads the actual stock again.  The stamps
// and the samples use all the workgroup's threads, the engine's step loop its first
// wavefront.  Nothing returns to the host between sub-steps; many programs share one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mcsim.h"
#include "engine_dev.h"
#include "mc_host.h"
#include "mc_math.h"

// the library's engine handle (engine.hip); its first member is the host mirror
struct mc_engine;

using namespace mce;

namespace {

struct StockDev {
  float *H, *err, *wallD, *markS, *markF;
  int32_t *lineCut, *lineWall, *rect;
  float* smax;   // exact maxima of the SW x SH node strips, snx per row of strips, sny rows
  int32_t nx, ny, snx, sny;
  double hx, hy, zTop, zB, band;
};

struct ProgDev {
  EngDev* eng;
  StockDev act;
  const mc_substep* sub;
  mc_substep_out* out;
  const float* lut;   // bull-nose: the tip table (null otherwise)
  mc_frame_out* frames;
  double R, fluteLen, rw, share, rc, dynBudget, fz0;
  int32_t nsub, ball, N, lutn, nframes;
  int32_t hOnly;      // flat or ball end: the actual stamps write the heights alone (the rest after the launch)
  // the thin-wall model (flex): the half-cell walk table, its entries up to 20 and 60 mm, the
  // step, the material and the tool's static stiffness, and the per-sub-step records
  const double* sacc;
  mc_wall_out* walls;
  double wstep, wallE, wallRho, kTool;
  int32_t flex, n20, n60, pad2;
  EdgeProfile edge;   // an imported edge profile (edge.r null otherwise)
  double dynLeft0;    // the frame budget left when the call starts (a call may begin inside a frame)
  double* dynOut;     // the budget left after the last sub-step, or null
};

// Stock.sample (src/stock.js:85-92)
__device__ __forceinline__ double sample_at(const StockDev& S, double x, double y) {
  const double fx = x / S.hx, fy = y / S.hy;
  if (fx < 0 || fy < 0 || fx > S.nx - 1 || fy > S.ny - 1) return -__builtin_inf();
  int i = (int)fx, j = (int)fy;
  if (i > S.nx - 2) i = S.nx - 2;
  if (j > S.ny - 2) j = S.ny - 2;
  const double tx = fx - i, ty = fy - j;
  const float* H = S.H;
  const double a = H[j * S.nx + i], b = H[j * S.nx + i + 1], c = H[(j + 1) * S.nx + i],
               d = H[(j + 1) * S.nx + i + 1];
  return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}

constexpr int NT = 256;   // threads per program: the stamps and samples use all, the engine wave 0

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = jmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// block-wide max / sum through LDS (red holds NT / 64 doubles); every thread gets the result
__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < NT / 64; w++) m = jmax(m, red[w]);
  return m;
}
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = 0;
  for (int w = 0; w < NT / 64; w++) m += red[w];
  return m;
}

// the node box of Stock.stamp (src/stock.js:119-128): empty when the tool is above the top
struct Box { int i0, i1, j0, j1; bool ok; };
__device__ __forceinline__ Box stamp_box(const StockDev& S, double cx, double cy, double tipZ, double R) {
  Box b{0, -1, 0, -1, false};
  if (tipZ >= S.zTop) return b;
  const double Rw = R + jmax(2 * jmax(S.hx, S.hy), S.band);
  const double fi0 = jmax(0, __builtin_floor((cx - Rw) / S.hx)), fi1 = jmin(S.nx - 1, __builtin_ceil((cx + Rw) / S.hx));
  const double fj0 = jmax(0, __builtin_floor((cy - Rw) / S.hy)), fj1 = jmin(S.ny - 1, __builtin_ceil((cy + Rw) / S.hy));
  if (!(fi0 <= fi1) || !(fj0 <= fj1)) return b;
  b.i0 = (int)fi0; b.i1 = (int)fi1; b.j0 = (int)fj0; b.j1 = (int)fj1; b.ok = true;
  return b;
}

// Stock.stamp's node loop (src/stock.js:129-155), flat, ball or a tip table (LUT), by
// threads t = 0..nt-1 of the workgroup: the nodes are independent, so they are taken in any
// order, U at a time with their loads issued before any store.  Returns the thread's part
// of the removed height sum; a node changed sets *touched.
template <bool LUT>
__device__ __forceinline__ double stamp_nodes(const StockDev& S, const Box& bx, double cx, double cy, double tipZ,
                                              double R, bool ball, const float* lut, int lutn, double errVal,
                                              double mux, double muy, double mfeed, int32_t line, int t, int nt,
                                              int* touched) {
  const double Rw = R + jmax(2 * jmax(S.hx, S.hy), S.band), Rw2 = Rw * Rw;
  const double R2 = R * R;
  const double mf = mfeed > 0 && !ball && !LUT ? mfeed : 0;
  const double zRim = tipZ + (LUT ? (double)lut[lutn - 1] : ball ? R : 0);
  const int W = bx.i1 - bx.i0 + 1, total = W * (bx.j1 - bx.j0 + 1);
  constexpr int U = 8;
  double vol = 0;
  bool tch = false;
  for (int q0 = t; q0 < total; q0 += U * nt) {
    int64_t k[U];
    float hv[U], wd[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = q0 + u * nt;
      const int qq = q < total ? q : q0;
      const int j = bx.j0 + qq / W, i = bx.i0 + qq % W;
      k[u] = (int64_t)j * S.nx + i;
      hv[u] = S.H[k[u]];
      wd[u] = S.wallD[k[u]];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = q0 + u * nt;
      if (q >= total) continue;
      const int j = bx.j0 + q / W, i = bx.i0 + q % W;
      const double dy = j * S.hy - cy, dx = i * S.hx - cx;
      const double r2 = dx * dx + dy * dy;
      if (r2 >= R2 - 0.003 * R) {
        if (r2 <= Rw2 && (double)hv[u] > zRim) {
          const double d = __builtin_sqrt(r2) - R;
          if (d < (double)wd[u]) { S.wallD[k[u]] = (float)d; S.lineWall[k[u]] = line; tch = true; }
        }
        continue;
      }
      const double z = LUT ? tipZ + (double)lut[(int)jmin(lutn - 1, mcm::round_js(__builtin_sqrt(r2) / MC_LUT_DR))]
                       : ball ? tipZ + R - __builtin_sqrt(R2 - r2) : tipZ;
      double h = hv[u];
      if (h > z) {
        vol += h - z;
        const float nh = (float)jmax(z, S.zB);
        S.H[k[u]] = nh;
        S.err[k[u]] = (float)errVal;
        S.wallD[k[u]] = __builtin_inff();
        S.lineCut[k[u]] = line;
        tch = true;
        h = nh;
      }
      if (mf != 0 && h > z - 0.002) {
        const double xs = i * S.hx, ys = j * S.hy, p = -dx * muy + dy * mux;
        S.markS[k[u]] = (float)(xs * mux + ys * muy + __builtin_sqrt(jmax(0, R2 - p * p)));
        S.markF[k[u]] = (float)mf;
        tch = true;
      }
    }
  }
  if (tch) *touched = 1;
  return vol;
}

// the strips the stamps skip by their maxima: SW x SH nodes, a wavefront's 64 lanes, whose
// loads cover two rows of 128 contiguous bytes
constexpr int SW = mch::MCH_STRIP_W, SH = mch::MCH_STRIP_H;
static_assert(SW * SH == 64, "a strip is one wavefront");

// each strip's highest node, one wavefront per strip (4 per workgroup)
__global__ __launch_bounds__(256) void strip_max(const float* __restrict__ H, int nx, int ny, int snx, int sny,
                                                 float* __restrict__ smax) {
  const int strip = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (strip >= snx * sny) return;
  const int si = strip % snx, sj = strip / snx;
  const int i = si * SW + (lane & (SW - 1)), j = sj * SH + lane / SW;
  float h = i < nx && j < ny ? H[(int64_t)j * nx + i] : -__builtin_inff();
  for (int o = 32; o > 0; o >>= 1) h = fmaxf(h, __shfl_xor(h, o, 64));
  if (lane == 0) smax[strip] = h;
}

// Stock.stamp's node loop (src/stock.js:129-155), flat, ball or a tip table (LUT), over the
// SW x SH node strips that meet the box, one strip per wavefront at a time with a lane per
// node (waves w = 0..nw-1 of the workgroup take strips w, w + nw, ...).
// Every update of a stamp needs the node above the stamp's lowest reach over it: the cut
// H > z, the wall distance H > zRim >= z, the feed mark H > tipZ - 0.002.  So a strip whose
// maximum is at or below a lower bound of that reach over the block is skipped, exactly:
// tipZ for a flat end (tipZ - 0.002 with feed marks), tipZ + min(lut) for a tip table, and
// for a ball the sphere's height at the block's point nearest the stamp's axis, less a
// margin far above the rounding of the per-node height; strips beyond the wall band are
// skipped too.  The nodes are independent, so the strips go in any order; a strip with a
// lowered node gets its maximum again from its nodes, so the maxima stay exact (heights
// only fall).  Returns the thread's part of the removed height sum; a node changed sets
// *touched.
template <bool LUT>
__device__ __forceinline__ double stamp_blocks(const StockDev& S, const Box& bx, double cx, double cy, double tipZ,
                                               double R, bool ball, const float* lut, int lutn, double lutMin,
                                               double errVal, double mux, double muy, double mfeed, int32_t line,
                                               int wave, int nw, int* touched) {
  const int lane = threadIdx.x & 63;
  const double Rw = R + jmax(2 * jmax(S.hx, S.hy), S.band), Rw2 = Rw * Rw;
  const double R2 = R * R, tie = R2 - 0.003 * R;
  const double mf = mfeed > 0 && !ball && !LUT ? mfeed : 0;
  const double zRim = tipZ + (LUT ? (double)lut[lutn - 1] : ball ? R : 0);
  // the stamp's lowest reach where it is uniform over the stamp
  const double zflat = LUT ? tipZ + lutMin : mf != 0 ? tipZ - 0.002 : tipZ;
  const int bi0 = bx.i0 / SW, bi1 = bx.i1 / SW, bj0 = bx.j0 / SH, bj1 = bx.j1 / SH;
  const int bw = bi1 - bi0 + 1, nb = bw * (bj1 - bj0 + 1);
  double vol = 0;
  bool tch = false;
  for (int b0 = wave * 64; b0 < nb; b0 += nw * 64) {
    // lane l tests block b0 + l: its maximum against the stamp's lowest reach over it
    bool live = false;
    {
      const int q = b0 + lane;
      if (q < nb) {
        const int bi = bi0 + q % bw, bj = bj0 + q / bw;
        // the block's part of the box, and its nearest point to the stamp's axis (a lower
        // bound of every node's r2, less a margin above the rounding of r2)
        const int ia = bi * SW > bx.i0 ? bi * SW : bx.i0, ib = bi * SW + SW - 1 < bx.i1 ? bi * SW + SW - 1 : bx.i1;
        const int ja = bj * SH > bx.j0 ? bj * SH : bx.j0, jb = bj * SH + SH - 1 < bx.j1 ? bj * SH + SH - 1 : bx.j1;
        const double ex = jmax(0, jmax(ia * S.hx - cx, cx - ib * S.hx));
        const double ey = jmax(0, jmax(ja * S.hy - cy, cy - jb * S.hy));
        const double rmin2 = (ex * ex + ey * ey) * (1 - 1e-12);
        // beyond the wall band nothing changes; past the cutting radius only the wall
        // distance, which needs H > zRim; inside it the cut (and marks) need H above the
        // stamp's lowest reach over the block
        const double zlow = rmin2 >= tie ? zRim
                          : ball && !LUT ? tipZ + R - __builtin_sqrt(R2 - rmin2) - 1e-9 : zflat;
        live = rmin2 <= Rw2 && (double)S.smax[bj * S.snx + bi] > zlow;
      }
    }
    unsigned long long m = __ballot(live);
    // the live blocks, eight at a time: their nodes' loads issued before any store
    while (m) {
      constexpr int U = 8;
      int qb[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        qb[u] = m ? b0 + __builtin_ctzll(m) : -1;
        m &= m ? m - 1 : 0ull;
      }
      int64_t k[U];
      int ni[U], nj[U];
      bool inS[U], inB[U];
      double h[U];
      float wd[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int q = qb[u] >= 0 ? qb[u] : qb[0];
        const int bi = bi0 + q % bw, bj = bj0 + q / bw;
        ni[u] = bi * SW + (lane & (SW - 1)); nj[u] = bj * SH + lane / SW;
        inS[u] = qb[u] >= 0 && ni[u] < S.nx && nj[u] < S.ny;
        inB[u] = inS[u] && ni[u] >= bx.i0 && ni[u] <= bx.i1 && nj[u] >= bx.j0 && nj[u] <= bx.j1;
        // every lane loads (a node clamped into the stock where its own is outside), so that the
        // loads issue as one batch; the values of lanes outside are not used
        k[u] = (int64_t)(nj[u] < S.ny ? nj[u] : S.ny - 1) * S.nx + (ni[u] < S.nx ? ni[u] : S.nx - 1);
        const float hv = S.H[k[u]];
        wd[u] = S.wallD[k[u]];
        h[u] = inS[u] ? (double)hv : -__builtin_inf();
      }
      bool cutb[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        cutb[u] = false;
        if (qb[u] < 0) continue;   // wave-uniform
        const int i = ni[u], j = nj[u];
        bool cut = false;
        if (inB[u]) {
          const double dy = j * S.hy - cy, dx = i * S.hx - cx;
          const double r2 = dx * dx + dy * dy;
          if (r2 >= tie) {
            if (r2 <= Rw2 && h[u] > zRim) {
              const double d = __builtin_sqrt(r2) - R;
              if (d < (double)wd[u]) { S.wallD[k[u]] = (float)d; S.lineWall[k[u]] = line; tch = true; }
            }
          } else {
            const double z = LUT ? tipZ + (double)lut[(int)jmin(lutn - 1, mcm::round_js(__builtin_sqrt(r2) / MC_LUT_DR))]
                           : ball ? tipZ + R - __builtin_sqrt(R2 - r2) : tipZ;
            if (h[u] > z) {
              vol += h[u] - z;
              const float nh = (float)jmax(z, S.zB);
              S.H[k[u]] = nh;
              S.err[k[u]] = (float)errVal;
              S.wallD[k[u]] = __builtin_inff();
              S.lineCut[k[u]] = line;
              tch = true;
              cut = true;
              h[u] = nh;
            }
            if (mf != 0 && h[u] > z - 0.002) {
              const double xs = i * S.hx, ys = j * S.hy, p = -dx * muy + dy * mux;
              S.markS[k[u]] = (float)(xs * mux + ys * muy + __builtin_sqrt(jmax(0, R2 - p * p)));
              S.markF[k[u]] = (float)mf;
              tch = true;
            }
          }
        }
        cutb[u] = __any(cut);
      }
      // a block with a lowered node gets its maximum again from its nodes (f32 heights: max is
      // exact in any order), the blocks' reductions interleaved; heights only fall, so an
      // untouched block keeps its maximum
      float hm[U];
#pragma unroll
      for (int u = 0; u < U; u++) hm[u] = (float)h[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int u = 0; u < U; u++) hm[u] = fmaxf(hm[u], __shfl_xor(hm[u], o, 64));
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (cutb[u] && lane == 0) S.smax[(nj[u] / SH) * S.snx + ni[u] / SW] = hm[u];
    }
  }
  if (tch) *touched = 1;
  return vol;
}

// A flat or ball end's stamp on the heights alone, for the chain: each sub-step reads only
// the heights the previous one left, so the deflected stamps' errors, wall distances, lines,
// feed marks, volumes and dirty box are written after the launch by one batch with
// Stock.stamp's sequential semantics on the heights saved before it.  A node inside the
// cutting radius with H > z (tipZ, or the sphere's height over the node) takes
// f32(max(z, zB)) (src/stock.js:141-147).  Flat: stock heights are never below f32(zB) and no
// f32 value lies strictly between tipZ and its rounding nh, so a node with H at or below
// max(tipZ, nh) keeps its height, and a strip whose maximum is at or below that is skipped
// exactly (the floor the previous sub-step left at this depth).  Ball: a strip whose maximum
// lies at or below the sphere's height at its point nearest the axis (less a margin above the
// rounding) has no node above the sphere.  A strip with a lowered node gets its maximum again
// from its nodes.
__device__ __forceinline__ void stamp_heights(const StockDev& S, const Box& bx, double cx, double cy, double tipZ,
                                              double R, bool ball, int wave, int nw) {
  const int lane = threadIdx.x & 63;
  const double R2 = R * R, tie = R2 - 0.003 * R;
  const float nh = (float)jmax(tipZ, S.zB);
  const double zskip = jmax(tipZ, (double)nh);
  const int bi0 = bx.i0 / SW, bi1 = bx.i1 / SW, bj0 = bx.j0 / SH, bj1 = bx.j1 / SH;
  const int bw = bi1 - bi0 + 1, nb = bw * (bj1 - bj0 + 1);
  for (int b0 = wave * 64; b0 < nb; b0 += nw * 64) {
    bool live = false;
    {
      const int q = b0 + lane;
      if (q < nb) {
        const int bi = bi0 + q % bw, bj = bj0 + q / bw;
        const int ia = bi * SW > bx.i0 ? bi * SW : bx.i0, ib = bi * SW + SW - 1 < bx.i1 ? bi * SW + SW - 1 : bx.i1;
        const int ja = bj * SH > bx.j0 ? bj * SH : bx.j0, jb = bj * SH + SH - 1 < bx.j1 ? bj * SH + SH - 1 : bx.j1;
        const double ex = jmax(0, jmax(ia * S.hx - cx, cx - ib * S.hx));
        const double ey = jmax(0, jmax(ja * S.hy - cy, cy - jb * S.hy));
        const double rmin2 = (ex * ex + ey * ey) * (1 - 1e-12);
        const double zlow = ball ? tipZ + R - __builtin_sqrt(R2 - rmin2) - 1e-9 : zskip;
        live = rmin2 < tie && (double)S.smax[bj * S.snx + bi] > zlow;
      }
    }
    unsigned long long m = __ballot(live);
    while (m) {
      constexpr int U = 8;
      int qb[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        qb[u] = m ? b0 + __builtin_ctzll(m) : -1;
        m &= m ? m - 1 : 0ull;
      }
      int64_t k[U];
      int ni[U], nj[U];
      bool inB[U];
      float h[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int q = qb[u] >= 0 ? qb[u] : qb[0];
        const int bi = bi0 + q % bw, bj = bj0 + q / bw;
        ni[u] = bi * SW + (lane & (SW - 1)); nj[u] = bj * SH + lane / SW;
        const bool inS = qb[u] >= 0 && ni[u] < S.nx && nj[u] < S.ny;
        inB[u] = inS && ni[u] >= bx.i0 && ni[u] <= bx.i1 && nj[u] >= bx.j0 && nj[u] <= bx.j1;
        k[u] = (int64_t)(nj[u] < S.ny ? nj[u] : S.ny - 1) * S.nx + (ni[u] < S.nx ? ni[u] : S.nx - 1);
        const float hv = S.H[k[u]];
        h[u] = inS ? hv : -__builtin_inff();
      }
      bool cutb[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        cutb[u] = false;
        if (qb[u] < 0) continue;   // wave-uniform
        bool cut = false;
        if (inB[u]) {
          const double dy = nj[u] * S.hy - cy, dx = ni[u] * S.hx - cx;
          const double r2 = dx * dx + dy * dy;
          if (r2 < tie) {
            const double z = ball ? tipZ + R - __builtin_sqrt(R2 - r2) : tipZ;
            if ((double)h[u] > z) {
              const float v = ball ? (float)jmax(z, S.zB) : nh;
              S.H[k[u]] = v;
              cut = true;
              h[u] = v;
            }
          }
        }
        cutb[u] = __any(cut);
      }
      float hm[U];
#pragma unroll
      for (int u = 0; u < U; u++) hm[u] = h[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int u = 0; u < U; u++) hm[u] = fmaxf(hm[u], __shfl_xor(hm[u], o, 64));
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (cutb[u] && lane == 0) S.smax[(nj[u] / SH) * S.snx + ni[u] / SW] = hm[u];
    }
  }
}

// The tip-table form runs in a code path of its own: mixed into the flat / ball loop it
// raised the kernel's register pressure into scratch spills.  A ball's stamps skip strips by
// their maxima (the floor a ball leaves lies above most of the next stamp's sphere); a flat
// end with all its maps rewrites every floor node under it (the feed marks, and the cut
// wherever the f32 floor height rounded above tipZ), so this form takes every node of the
// box (the pipeline's flat and ball programs take stamp_heights instead).
__device__ __forceinline__ double stamp_part(const StockDev& S, const Box& bx, double cx, double cy, double tipZ,
                                             double R, bool ball, const float* lut, int lutn, double errVal,
                                             double mux, double muy, double mfeed, int32_t line, int t, int nt,
                                             int* touched) {
  if (!bx.ok) return 0;
  if (lut) return stamp_nodes<true>(S, bx, cx, cy, tipZ, R, ball, lut, lutn, errVal, mux, muy, mfeed, line, t, nt, touched);
  if (ball)
    return stamp_blocks<false>(S, bx, cx, cy, tipZ, R, true, nullptr, 0, 0.0, errVal, mux, muy, mfeed, line, t >> 6,
                               nt >> 6, touched);
  return stamp_nodes<false>(S, bx, cx, cy, tipZ, R, ball, nullptr, 0, errVal, mux, muy, mfeed, line, t, nt, touched);
}

// ---- The thin-wall model (src/wall.js), read from the actual stock by the whole workgroup.
// Every walk of the reference steps by half a cell with a running sum (s += step); the sums
// from 0 are the table sacc (sacc[k] = s after k + 1 additions, up to 60 mm, made on the host
// in the same order), so the walks from 0 are searched in parallel over the table, and the
// walks from another start (the far floor from t, the thickness across from s0) are summed
// lane by lane, 64 elements per pass, each lane keeping its own.
constexpr int WALL_DIRS = 17;   // -40 .. 40 degrees in 5 degree steps (src/wall.js:26)
constexpr int WALL_NZ = 10;     // heights of the thickness profile (src/wall.js:92)
enum { W_SX, W_SY, W_N0, W_CX, W_CY, W_T, W_NX, W_NY, W_BASE, W_A, W_HW, W_MX, W_MY, W_COUNT };
// block-uniform decisions, one slot per phase, so that no thread overwrites a flag that a
// slower thread has still to read
enum { F_CONTACT, F_DIRS, F_BEST, F_MEASURE, F_ALONG, F_FOUND, F_COUNT };

struct WallLds {
  double tx[MC_BINS / 2], ty[MC_BINS / 2];   // the contact terms of the engaged bins
  double dcs[WALL_DIRS], dsn[WALL_DIRS];     // the crossing directions
  double thick[WALL_NZ + 1];                 // thicknessAt at the load height, then the profile
  double ts[WALL_NZ];                        // the profile scaled to the load height's thickness
  double v[W_COUNT];
  mc_wall_out rec;                           // the sub-step's record
  int dirk[WALL_DIRS];                       // first crossing per direction (table index)
  int lk[2];                                 // first break along the wall, each way
  int go[F_COUNT];
};

// Stock.floor (src/stock.js:94-97)
__device__ __forceinline__ double floor_at(const StockDev& S, double x, double y) {
  const double v = sample_at(S, x, y);
  return v == -__builtin_inf() ? S.zB : v;
}

// cellAt (src/wall.js:46-50): the nearest node's height
__device__ __forceinline__ double cell_at(const StockDev& S, double x, double y) {
  const double i = mcm::round_js(x / S.hx), j = mcm::round_js(y / S.hy);
  if (!(i >= 0) || !(j >= 0) || i >= S.nx || j >= S.ny) return -__builtin_inf();
  return S.H[(int64_t)j * S.nx + (int64_t)i];
}

__device__ __forceinline__ unsigned long long ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// thicknessAt (src/wall.js:54-60) by one wavefront: back out of the material from the
// contact point (s0 = 0, -sacc[0], -sacc[1], ... while s0 > -6 and the node is at or above
// z), then across from s0 + step while s <= s0 + 20.  Every lane returns the answer.
__device__ double thickness_at_wave(const StockDev& S, const double* sacc, int nsacc, double step, double cx,
                                    double cy, double nx, double ny, double z, int lane) {
  int m = 0;
  for (int base = 0;; base += 64) {   // wave-uniform
    const int k = base + lane;
    const double s0 = k == 0 ? 0.0 : k - 1 < nsacc ? -sacc[k - 1] : -__builtin_inf();
    const bool stop = !(s0 > -6) || !(cell_at(S, cx + nx * s0, cy + ny * s0) >= z);
    const unsigned long long b = ballot64(stop);
    if (b) { m = base + __builtin_ctzll(b); break; }
  }
  const double s0 = m == 0 ? 0.0 : -sacc[m - 1 < nsacc ? m - 1 : nsacc - 1], lim = s0 + 20;
  double cur = s0;   // the element before this pass's first
  for (;;) {         // wave-uniform: the elements grow, so those within lim come first
    double mine = 0;
    for (int j = 0; j < 64; j++) { cur = cur + step; mine = j == lane ? cur : mine; }
    const bool in = mine <= lim;
    const unsigned long long h = ballot64(in && cell_at(S, cx + nx * mine, cy + ny * mine) < z);
    if (h) return __shfl(mine, __builtin_ctzll(h), 64) - s0;
    if (ballot64(!in)) return -1;
  }
}

// wallAt (src/wall.js:11-34) with measure (:62-112), called by every thread of the
// workgroup; Hb is the engine's map, (ux, uy) the feed direction.  Returns whether a wall was
// found (the same on every thread); its fields are in W.rec.
__device__ bool wall_at_block(const StockDev& S, const float* Hb, const double* sacc, int n20, int n60, double step,
                              double px, double py, double pz, double R, double ux, double uy, double zLoad,
                              double E, double rho, double kTool, WallLds& W) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double PI = 3.141592653589793;
  // the contact point: the middle of the engaged arc at the load height (:12-23); a bin
  // left out adds +0, which leaves a sum started at +0 unchanged
  for (int b = tid; b < MC_BINS / 2; b += NT) {
    double tx = 0, ty = 0;
    if (!((double)Hb[b] <= zLoad)) {
      const double phi = (b + 0.5) / MC_BINS * 2 * PI;
      double sn, cs;
      mcm::sincos_js(phi, sn, cs);
      tx = sn * ux + cs * -uy;
      ty = sn * uy + cs * ux;
    }
    W.tx[b] = tx; W.ty[b] = ty;
  }
  __syncthreads();
  if (tid < 2) {   // the two sums in bin order, side by side on lanes 0 and 1
    const double* a = tid == 0 ? W.tx : W.ty;
    double s = 0;
    for (int b = 0; b < MC_BINS / 2; b++) s += a[b];
    W.v[W_SX + tid] = s;
  }
  __syncthreads();
  if (tid == 0) {
    const double sx = W.v[W_SX], sy = W.v[W_SY], L = mcm::hypot_js(sx, sy);
    W.go[F_CONTACT] = !(L < 1e-9);
    if (W.go[F_CONTACT]) {
      W.v[W_N0] = mcm::atan2_js(sy, sx);
      W.v[W_CX] = px + R * sx / L;
      W.v[W_CY] = py + R * sy / L;
    }
  }
  __syncthreads();
  if (!W.go[F_CONTACT]) return false;
  if (tid < WALL_DIRS) {
    const double d = -40 + 5 * tid, a = W.v[W_N0] + d * PI / 180;
    double sn, cs;
    mcm::sincos_js(a, sn, cs);
    W.dcs[tid] = cs; W.dsn[tid] = sn;
    W.dirk[tid] = 0x7fffffff;
  }
  __syncthreads();
  const double cx = W.v[W_CX], cy = W.v[W_CY];
  // thickness (:36-42) along every direction: the first sacc[k] <= 20 whose bilinear sample
  // lies below zLoad; 15 threads per direction, 8 steps each per round
  for (int base = 0; base < n20; base += 120) {   // block-uniform
    if (tid < 15 * WALL_DIRS) {
      const int i = tid / 15, o = tid - 15 * i;
      const double nx = W.dcs[i], ny = W.dsn[i];
      int hit = 0x7fffffff;
      for (int j = 7; j >= 0; j--) {
        const int k = base + o + 15 * j;
        if (k < n20) {
          const double s = sacc[k];
          if (sample_at(S, cx + nx * s, cy + ny * s) < zLoad) hit = k;
        }
      }
      if (hit != 0x7fffffff) atomicMin(&W.dirk[i], hit);
    }
    __syncthreads();
    if (tid == 0) {
      int all = 1;
      for (int i = 0; i < WALL_DIRS; i++) all &= W.dirk[i] != 0x7fffffff;
      W.go[F_DIRS] = all;
    }
    __syncthreads();
    if (W.go[F_DIRS]) break;
  }
  // the thinnest crossing, the first of equals in the loop's order (:27-31)
  if (tid == 0) {
    int bi = -1;
    double bt = 0;
    for (int i = 0; i < WALL_DIRS; i++) {
      if (W.dirk[i] == 0x7fffffff) continue;
      const double t = sacc[W.dirk[i]];
      if (t > 0 && (bi < 0 || t < bt)) { bi = i; bt = t; }
    }
    W.go[F_BEST] = bi >= 0;
    if (bi >= 0) { W.v[W_T] = bt; W.v[W_NX] = W.dcs[bi]; W.v[W_NY] = W.dsn[bi]; }
  }
  __syncthreads();
  if (!W.go[F_BEST]) return false;
  const double t = W.v[W_T], nx = W.v[W_NX], ny = W.v[W_NY];
  // measure (:62-72): the far-side floor over [t, t + 3], the wall's base, free height and
  // load height; slivers are left out
  if (wv == 0) {
    double cur = t, fm = __builtin_inf();
    const double lim = t + 3;
    for (;;) {   // wave-uniform
      double mine = 0;
      for (int j = 0; j < 64; j++) { mine = j == lane ? cur : mine; cur = cur + step; }
      const bool in = mine <= lim;
      if (in) fm = jmin(fm, floor_at(S, cx + nx * mine, cy + ny * mine));
      if (ballot64(!in)) break;
    }
    for (int o = 32; o > 0; o >>= 1) fm = jmin(fm, __shfl_xor(fm, o, 64));
    if (lane == 0) {
      const double base = jmax(pz, fm), mx = cx + nx * t / 2, my = cy + ny * t / 2;
      const double hWall = sample_at(S, mx, my) - base;
      W.v[W_BASE] = base; W.v[W_MX] = mx; W.v[W_MY] = my; W.v[W_HW] = hWall;
      W.v[W_A] = jmin(jmax(zLoad - base, 0.3), jmax(hWall, 0.3));
      W.go[F_MEASURE] = !(hWall < 0.5 || t < 2 * step * 2);
      W.lk[0] = 0x7fffffff; W.lk[1] = 0x7fffffff;
    }
  }
  __syncthreads();
  if (!W.go[F_MEASURE]) return false;
  // the length along the wall at the load height (:74-83): the first break each way, the
  // steps from 0 up to 60 mm read from the table
  const double mx = W.v[W_MX], my = W.v[W_MY], tx = -ny, ty = nx;
  for (int base = 0; base < n60; base += 4 * (NT / 2)) {   // block-uniform
    const int side = tid / (NT / 2), o = tid - side * (NT / 2);
    const double sg = side == 0 ? 1.0 : -1.0;
    int hit = 0x7fffffff;
    for (int j = 3; j >= 0; j--) {
      const int k = base + o + (NT / 2) * j;
      if (k < n60) {
        const double s = sacc[k];
        if (sample_at(S, mx + tx * s * sg, my + ty * s * sg) < zLoad) hit = k;
      }
    }
    if (hit != 0x7fffffff) atomicMin(&W.lk[side], hit);
    __syncthreads();
    if (tid == 0) W.go[F_ALONG] = W.lk[0] != 0x7fffffff && W.lk[1] != 0x7fffffff;
    __syncthreads();
    if (W.go[F_ALONG]) break;
  }
  const double base = W.v[W_BASE], a = W.v[W_A], hWall = W.v[W_HW];
  // the thickness profile (:88-96): thicknessAt at the load height and at ten heights from
  // the root, three calls per wavefront
  for (int c = wv; c < WALL_NZ + 1; c += NT / 64) {   // wave-uniform
    const double z = c == 0 ? zLoad : base + (c - 1 + 0.5) / WALL_NZ * a;
    const double tq = thickness_at_wave(S, sacc, n60, step, cx, cy, nx, ny, z, lane);
    if (lane == 0) W.thick[c] = tq;
  }
  __syncthreads();
  if (tid == 0) {
    // the length: m1 steps one way (the table's sum), then m2 more steps the other way, summed
    // on; once it reaches both 2 t and the plate-width bound nothing it decides can change
    const int m1 = W.lk[0] == 0x7fffffff ? n60 : W.lk[0], m2 = W.lk[1] == 0x7fffffff ? n60 : W.lk[1];
    const double X = 2 * a + t + R;
    double L = m1 > 0 ? sacc[m1 - 1] : 0.0;
    for (int j = 0; j < m2 && !(L >= X && L >= 2 * t); j++) L += step;
    bool ok = !(L < 2 * t);
    if (ok) {
      // the cantilever strip's compliance at the load height (:97-104) and its mass (:105-110)
      const double bEff = jmin(L, X), tRef = W.thick[0];
      double* ts = W.ts;
      double comp = 0, tMean = 0;
      for (int q = 0; q < WALL_NZ; q++) {
        const double tq = W.thick[1 + q];
        ts[q] = tq > 0 && tRef > 0 ? jmax(t * tq / tRef, t) : t;
      }
      for (int q = 0; q < WALL_NZ; q++) {
        const double z = (q + 0.5) / WALL_NZ * a, Iq = bEff * mcm::pow_js(ts[q], 3.0) / 12;
        comp += (a - z) * (a - z) / (E * Iq) * (a / WALL_NZ);
      }
      const double k = 1 / comp, I = bEff * mcm::pow_js(t, 3.0) / 12;
      ok = !(k > 20 * kTool);
      if (ok) {
        for (int q = 0; q < WALL_NZ; q++) tMean = tMean + ts[q];
        tMean = tMean / WALL_NZ;
        double m = 0.24 * rho * 1e-9 * tMean * bEff * hWall;
        double fn = __builtin_sqrt(k * 1e3 / m) / (2 * PI);
        if (fn > MC_WALL_FN_CAP) {
          const double wc = 2 * PI * MC_WALL_FN_CAP;
          fn = MC_WALL_FN_CAP;
          m = k * 1e3 / (wc * wc);
        }
        mc_wall_out& r = W.rec;
        r.k = k; r.m = m; r.fn = fn; r.zeta = 0.02; r.nx = nx; r.ny = ny; r.t = t; r.tRoot = ts[0];
        r.hWall = hWall; r.a = a; r.bEff = bEff; r.base = base; r.cx = cx; r.cy = cy; r.I = I;
        r.found = 1;
      }
    }
    W.go[F_FOUND] = ok;
  }
  __syncthreads();
  return W.go[F_FOUND] != 0;
}

// the dirty rectangle joins a changed stamp's box (src/stock.js:156-160)
__device__ __forceinline__ void join_rect(const StockDev& S, const Box& b) {
  atomicMin(&S.rect[0], b.i0); atomicMin(&S.rect[1], b.j0);
  atomicMax(&S.rect[2], b.i1); atomicMax(&S.rect[3], b.j1);
}

template <bool FLEX>
__global__ __launch_bounds__(NT) void sim_run(const ProgDev* progs) {
  __shared__ EngLds L;
  __shared__ double red[NT / 64];
  __shared__ int flags[2];   // flags[1]: the actual stamp changed a node
  __shared__ WallLds WL[1];
  const ProgDev& P = progs[blockIdx.x];
  EngDev& E = *P.eng;
  const int tid = threadIdx.x;
  const double PI = 3.141592653589793;
  const bool ball = P.ball != 0;
  const double R = P.R;
  const ToolShape shape{P.ball, R, P.rc, P.lut, P.lutn};
  const double cell = jmax(P.act.hx, P.act.hy);
  double fzLast = P.fz0;   // this.fz: the last cutting sub-step's feed per tooth
  double dynLeft = P.dynLeft0;    // the frame's dynamics budget left (src/sim.js:360)
  int frame = 0;                  // frames ended so far
  bool held = false;              // this.wall is not null (thread 0's)
  for (int k = 0; k < P.nsub; k++) {
    const mc_substep s = P.sub[k];
    if (s.frame & MC_FRAME_FIRST) dynLeft = P.dynBudget;
    const double vx = -s.uy, vy = s.ux;
    // ---- fillEngagement (src/sim.js:301-319); a vertical move keeps the previous map
    double hmax = -__builtin_inf();
    if (!(s.dxy < 1e-6)) {
      const double r = R - 0.5 * cell, ax = s.x + s.ux * cell, ay = s.y + s.uy * cell;
      for (int b = tid; b < MC_BINS; b += NT) {
        const double phi = (b + 0.5) / MC_BINS * 2 * PI;
        double sn, cs;
        mcm::sincos_js(phi, sn, cs);
        const double h = sample_at(P.act, ax + r * (sn * s.ux + cs * vx), ay + r * (sn * s.uy + cs * vy));
        const float f = sn > 0.02 ? (float)h : -__builtin_inff();
        E.Hb[b] = f;
        hmax = jmax(hmax, (double)f);
      }
      hmax = block_max(hmax, red);
    }
    const double ap = jmin(P.fluteLen, jmax(0, hmax - s.z));
    // ---- fillEndEngagement (src/sim.js:324-341) while descending
    const bool haveEnd = s.fza > 0;
    double dmax = 0, rOut = 0;
    if (haveEnd) {
      rOut = R - 0.5 * cell;
      for (int q = tid; q < MC_END_BINS * MC_END_RINGS; q += NT) {
        const int b = q / MC_END_RINGS, j = q - b * MC_END_RINGS;
        const double phi = (b + 0.5) / MC_END_BINS * 2 * PI;
        double sn, cs;
        mcm::sincos_js(phi, sn, cs);
        const double ex = sn * s.ux + cs * vx, ey = sn * s.uy + cs * vy;
        const double r = (j + 0.5) / MC_END_RINGS * rOut;
        const double zEdge = s.z + eng_zat(shape, r);
        const double d = jmax(0, sample_at(P.act, s.x + r * ex, s.y + r * ey) - zEdge);
        E.Eb[q] = (float)d;
        dmax = jmax(dmax, d);
      }
      dmax = block_max(dmax, red);
    }
    const bool endCut = haveEnd && dmax > 2e-4;
    const bool cut = (ap > 0.01 || endCut) && !s.rapid;
    // the engagement arc the simulator shows (src/sim.js:421-423): immersion(tipZ) over the
    // leading bins of the map (:343-348) when the flutes cut, [0, pi] for an end cut alone
    double phiSt = 0, phiEx = 0;
    int arc = 0;
    if (cut && ap > 0.01) {
      const bool up = tid < MC_BINS / 2 && (double)E.Hb[tid] > s.z + 0.1;
      const double last = block_max(up ? (double)tid : -1.0, red);
      const double first = -block_max(up ? -(double)tid : -1e9, red);
      if (last >= 0) {
        phiSt = jmax(0, (first + 0.5) / MC_BINS * 2 * PI - PI / MC_BINS);
        phiEx = jmin(PI, (last + 0.5) / MC_BINS * 2 * PI + PI / MC_BINS);
        arc = 1;
      }
    } else if (cut && endCut) {
      phiEx = PI;
      arc = 1;
    }
    int n = 0;
    // the commanded stamp does not depend on the dynamics: waves 1.. stamp it while wave 0
    // runs the engine
    if (tid == 0) {
      flags[0] = 0; flags[1] = 0;
      if (FLEX) WL[0].rec = mc_wall_out{};
    }
    __threadfence_block();
    __syncthreads();
    if (cut) {
      // setEngagement, setEndEngagement (src/sim.js:403-404) on thread 0; the wall at the load
      // height with the thin-wall model, else none (:407), and setWork (:408)
      if (tid == 0) {
        eng_set_engagement(E, s.ux, s.uy, s.fzEff, s.z, ap, nullptr, nullptr, ball, P.rc, &P.edge);
        eng_set_end_engagement(E, endCut ? s.fza : 0, haveEnd ? rOut : 0, nullptr, shape);
      }
      bool found = false;
      if (FLEX && P.flex && ap > 0.01)   // block-uniform
        found = wall_at_block(P.act, E.Hb, P.sacc, P.n20, P.n60, P.wstep, s.x, s.y, s.z, R, s.ux, s.uy,
                              s.z + ap / 2, P.wallE, P.wallRho, P.kTool, WL[0]);
      if (tid == 0) {
        if (FLEX && found) {
          const mc_wall_out& w = WL[0].rec;
          eng_set_work(E, w.k, w.m, w.zeta, w.nx, w.ny);
        } else {
          eng_set_work_null(E);
        }
        held = found;
        E.hMax = jmax(0.1, 3 * E.fz);
      }
      __threadfence_block();
      __syncthreads();
      // the sub-step's dynamics: its own time, capped by the frame budget's share (:409-414)
      const double want = jmax(1, mcm::round_js(s.dtm / E.dt));
      if (P.dynBudget > 0) {
        const double cost = (double)E.N * jmax(1, E.nz + (endCut ? MC_END_RINGS : 0));
        const double share = jmax(E.spt / 4.0, dynLeft / cost / jmax(1, s.nSubEst - s.guard + 1));
        n = (int)jmax(1, jmin(want, __builtin_floor(share)));
        dynLeft -= n * cost;
      } else {
        n = (int)jmax(1, jmin(want, __builtin_floor(P.share)));
      }
      if (tid < 64) engine_steps(E, n, L);
      fzLast = s.fzEff;
    } else {
      // out of the cut the vibration decays, and a wall whose deflection has died out is
      // dropped (:427-429)
      if (tid == 0) {
        E.x *= 0.5; E.y *= 0.5; E.vx *= 0.5; E.vy *= 0.5; E.w *= 0.5; E.wv *= 0.5;
        if (truthy(E.wk) && __builtin_fabs(E.w) < 1e-7) { held = false; eng_set_work_null(E); }
      }
    }
    __threadfence_block();
    __syncthreads();
    // ---- deflection, then the actual stamp (:433-440)
    const double dx = (E.x - E.w * E.wnx) * 1e3, dy = (E.y - E.w * E.wny) * 1e3;
    const double errVal = mcm::hypot_js(E.x, E.y) * 1e6;
    const double maxF = cut ? E.runMaxF : 0.0;
    const double ax = s.x + dx, ay = s.y + dy, az = s.z + (ball ? P.rw : 0), aR = R - P.rw;
    const Box ba = stamp_box(P.act, ax, ay, az, aR);
    double vol = 0;   // with hOnly, the batch after the launch gives the volume
    if (P.hOnly) {   // block-uniform
      if (ba.ok) stamp_heights(P.act, ba, ax, ay, az, aR, ball, tid >> 6, NT / 64);
    } else {
      const double part = stamp_part(P.act, ba, ax, ay, az, aR, ball, P.lut, P.lutn, errVal, s.ux, s.uy,
                                     fzLast * P.N, s.line, tid, NT, &flags[1]);
      vol = block_sum(part, red) * P.act.hx * P.act.hy;
      if (tid == 0 && flags[1]) join_rect(P.act, ba);
    }
    __threadfence_block();
    __syncthreads();
    // the wall-error mark while a wall is held (:443): recorded for the caller's map
    if (FLEX && P.walls && tid == 0) {
      mc_wall_out& w = WL[0].rec;
      if (held && truthy(E.wk)) {
        const double v = E.runMeanW * 1e6;
        w.marked = 1; w.markVal = v > 0 ? v : 0.0; w.wnx = E.wnx; w.wny = E.wny;
      }
      P.walls[k] = w;
    }
    if (P.out && tid == 0) {
      mc_substep_out& o = P.out[k];
      o.dx = dx; o.dy = dy; o.ap = ap; o.errVal = errVal; o.vol = vol; o.maxF = maxF; o.nsteps = n; o.cut = cut ? 1 : 0;
      o.phiSt = phiSt; o.phiEx = phiEx; o.arc = arc; o.pad = 0;
    }
    // the frame's end (Simulator.collect, src/sim.js:644-663): takeStats() into the frame's
    // record, then chatterIndex(); the next sub-step's run starts from fresh stats
    if ((s.frame & MC_FRAME_LAST) && P.frames && frame < P.nframes) {
      if (tid == 0) {
        mc_frame_out& fo = P.frames[frame];
        double* st = (double*)&fo.stats;
        for (int q = 0; q < S_COUNT; q++) st[q] = E.st[q];
        eng_new_stats(E.st);
        double ci[3];
        eng_chatter_index(E, ci);
        fo.index = ci[0]; fo.amp = ci[1]; fo.once = ci[2];
        fo.end = k + 1; fo.pad = 0;
      }
      frame++;
      __threadfence_block();
      __syncthreads();
    }
  }
  if (P.dynOut && tid == 0) *P.dynOut = dynLeft;
}

}  // namespace

// engine.hip's accessors of a handle's device record and host mirror
EngDev* mce_engine_device(mc_engine* e);
EngDev* mce_engine_mirror(mc_engine* e);
int mce_engine_prepare(mc_engine* e, int64_t nsteps);
int mce_engine_finish(mc_engine* e);

extern "C" int mc_sim_run(const mc_sim_program* progs, int n) {
  if (n < 0 || (n > 0 && !progs)) return mch::fail(MC_ERR_ARG, "null programs or negative count");
  if (n == 0) return MC_OK;
  std::vector<ProgDev> pd((size_t)n);
  std::vector<const void*> seen;
  for (int p = 0; p < n; p++) {
    const mc_sim_program& g = progs[p];
    if (!g.actual || !g.ideal || !g.engine || g.actual == g.ideal)
      return mch::fail(MC_ERR_ARG, "program %d: needs an engine and two distinct stocks", p);
    if (g.nsub < 0 || (g.nsub > 0 && !g.sub)) return mch::fail(MC_ERR_ARG, "program %d: null sub-steps", p);
    if (g.type != MC_TOOL_FLAT && g.type != MC_TOOL_BALL && g.type != MC_TOOL_BULL && g.type != MC_TOOL_PROFILE)
      return mch::fail(MC_ERR_ARG, "program %d: flat, ball, bull-nose or profile tools", p);
    const bool tabled = g.type == MC_TOOL_BULL || g.type == MC_TOOL_PROFILE;
    if (g.type == MC_TOOL_PROFILE) {
      if (!g.edge_r || g.edge_n < 2 || g.edge_n > (1 << 20) || !(g.edge_dz > 0) || !std::isfinite(g.edge_dz))
        return mch::fail(MC_ERR_ARG, "program %d: a profile tool needs at least two edge radii and a positive step", p);
      for (int i = 0; i < g.edge_n; i++)
        if (!std::isfinite(g.edge_r[i])) return mch::fail(MC_ERR_ARG, "program %d: non-finite edge radius", p);
    }
    if (tabled) {
      if (!g.lut || g.lutn < 1 || g.lutn > (1 << 20) || !(g.cornerR >= 0))
        return mch::fail(MC_ERR_ARG, "program %d: a bull-nose or profile tool needs its tip table (and corner radius)", p);
      for (int i = 0; i < g.lutn; i++)
        if (!std::isfinite(g.lut[i])) return mch::fail(MC_ERR_ARG, "program %d: non-finite tip table", p);
    }
    if (!(g.R > 0) || !(g.rw >= 0) || !(g.rw < g.R) || !(g.fluteLen > 0) || !std::isfinite(g.fz) ||
        !(g.dynBudget > 0 ? std::isfinite(g.dynBudget) : g.share >= 1))
      return mch::fail(MC_ERR_ARG, "program %d: invalid radius, wear, flute length, feed, share or budget", p);
    if (g.nframes < 0 || (g.nframes > 0 && !g.frames))
      return mch::fail(MC_ERR_ARG, "program %d: null frames with a frame capacity", p);
    if (g.flex != 0 && g.flex != 1) return mch::fail(MC_ERR_ARG, "program %d: flex must be 0 or 1", p);
    // The device form of the thin-wall model (sim_run<true>, wall_at_block) is switched off:
    // its first run on an MI355X ended in a memory access fault whose cause is not found (the
    // same code run on the host, every buffer against a guard page, matches the oracle bit for
    // bit).  Until the cause is shown a program with flex is refused here, before any launch.
    if (g.flex)
      return mch::fail(MC_ERR_LIMIT, "program %d: the thin-wall model (flex) is switched off in the device pipeline", p);
    if (g.flex && !(std::isfinite(g.wallE) && g.wallE > 0 && std::isfinite(g.wallRho) && g.wallRho > 0 &&
                    std::isfinite(g.kTool) && g.kTool > 0))
      return mch::fail(MC_ERR_ARG, "program %d: the wall model needs a positive E, density and tool stiffness", p);
    for (const void* q : {(const void*)g.actual, (const void*)g.ideal, (const void*)g.engine})
      for (const void* o : seen)
        if (o == q) return mch::fail(MC_ERR_ARG, "program %d shares a stock or an engine with another", p);
    seen.push_back(g.actual); seen.push_back(g.ideal); seen.push_back(g.engine);
    for (int k = 0; k < g.nsub; k++) {
      const mc_substep& s = g.sub[k];
      if (!std::isfinite(s.x) || !std::isfinite(s.y) || !std::isfinite(s.z) || !std::isfinite(s.ux) ||
          !std::isfinite(s.uy) || !(s.dtm >= 0) || !std::isfinite(s.fzEff) || !std::isfinite(s.fza) || !(s.dxy >= 0) ||
          std::isnan(s.nSubEst) || s.frame < 0 || s.frame > (MC_FRAME_FIRST | MC_FRAME_LAST))
        return mch::fail(MC_ERR_ARG, "program %d sub-step %d: non-finite input or bad frame flags", p, k);
    }
    EngDev* m = mce_engine_mirror(g.engine);
    if (!m) return mch::fail(MC_ERR_ARG, "program %d: the engine is not configured", p);
    if (m->N * (MC_MAX_SLICES + MC_END_RINGS) > MC_MAX_FLUTES * (MC_MAX_SLICES + MC_END_RINGS))
      return mch::fail(MC_ERR_LIMIT, "program %d: too many flutes", p);
    ProgDev& d = pd[(size_t)p];
    mch::StockView a, i;
    if (int st = mch::stock_view(g.actual, &a)) return st;
    if (int st = mch::stock_view(g.ideal, &i)) return st;
    if (a.nx < 2 || a.ny < 2 || i.nx < 2 || i.ny < 2) return mch::fail(MC_ERR_ARG, "program %d: stocks need 2 x 2 nodes", p);
    float* sa = nullptr;
    int snxa = 0, snya = 0;
    if (g.type == MC_TOOL_BALL || g.type == MC_TOOL_FLAT)
      if (int st = mch::stock_strips(g.actual, &sa, &snxa, &snya)) return st;
    d.act = StockDev{a.H, a.err, a.wallD, a.markS, a.markF, a.lineCut, a.lineWall, a.rect, sa, a.nx, a.ny, snxa, snya,
                     a.hx, a.hy, a.zTop, a.zBottom, a.band};
    d.sub = nullptr; d.out = nullptr;
    d.R = g.R; d.fluteLen = g.fluteLen; d.rw = g.rw; d.share = g.share;
    d.nsub = g.nsub; d.ball = g.type == MC_TOOL_BALL; d.N = m->N;
    d.lut = nullptr; d.lutn = tabled ? g.lutn : 0;
    d.rc = g.type == MC_TOOL_BULL ? std::min(g.R, g.cornerR) : 0.0;
    d.edge = EdgeProfile{nullptr, g.type == MC_TOOL_PROFILE ? g.edge_n : 0, g.edge_dz};
    d.dynBudget = g.dynBudget > 0 ? g.dynBudget : 0; d.fz0 = g.fz;
    d.dynLeft0 = g.budgetLeft ? *g.budgetLeft : d.dynBudget;
    d.dynOut = nullptr;
    if (g.budgetLeft && !std::isfinite(d.dynLeft0))
      return mch::fail(MC_ERR_ARG, "program %d: non-finite frame budget left", p);
    d.frames = nullptr; d.nframes = g.frames ? g.nframes : 0;
    d.hOnly = g.type == MC_TOOL_FLAT || g.type == MC_TOOL_BALL ? 1 : 0;
    d.sacc = nullptr; d.walls = nullptr; d.flex = g.flex; d.n20 = 0; d.n60 = 0; d.pad2 = 0;
    d.wallE = g.wallE; d.wallRho = g.wallRho; d.kTool = g.kTool;
    // the wall model's walks step by half a cell (src/wall.js:37): at least 1e-4 mm, so that
    // a walk to 60 mm stays under 600,000 steps, and at most 5 mm, so that every walk has
    // steps before its end
    d.wstep = std::min(a.hx, a.hy) * 0.5;
    if (g.flex && !(d.wstep >= 1e-4 && d.wstep <= 5))
      return mch::fail(MC_ERR_LIMIT, "program %d: the wall model needs cells of 0.0002 to 10 mm", p);
  }
  // The commanded stamps (Simulator.ideal, src/sim.js:436) depend only on the schedule and
  // nothing in the chain reads the commanded stock, so each program's are applied as one
  // batch with Stock.stamp's sequential semantics (mc_stamp_batch_ex) instead of inside the
  // chain: R, the tool's type, the sub-step's position and line, errVal 0, no feed marks, and
  // a bull-nose tool's tip table.
  for (int p = 0; p < n; p++) {
    const mc_sim_program& g = progs[p];
    if (g.nsub == 0) continue;
    std::vector<mc_stamp> cmd((size_t)g.nsub);
    std::vector<mc_stamp_opt> opt(g.type == MC_TOOL_BULL || g.type == MC_TOOL_PROFILE ? (size_t)g.nsub : 0);
    for (int k = 0; k < g.nsub; k++) {
      const mc_substep& s = g.sub[k];
      cmd[(size_t)k] = mc_stamp{s.x, s.y, s.z, g.R, g.type == MC_TOOL_BALL ? MC_TOOL_BALL : MC_TOOL_FLAT, s.line, 0.0};
      if (!opt.empty()) opt[(size_t)k] = mc_stamp_opt{0.0, 0.0, 0.0, 0, 0};
    }
    const int32_t lo = 0, ln = g.lutn;
    const int st = opt.empty() ? mc_stamp_batch(g.ideal, cmd.data(), g.nsub, nullptr)
                               : mc_stamp_batch_ex(g.ideal, cmd.data(), opt.data(), g.nsub, g.lut, &lo, &ln, 1, nullptr);
    if (st) return st;
  }
  std::unique_lock<std::mutex> lk(mch::lock());
  hipStream_t q = mch::stream();
  // a ball's stamps skip node strips by their maxima and keep them exact: take them from
  // the heights first
  for (int p = 0; p < n; p++) {
    const StockDev& S = pd[(size_t)p].act;
    if (!S.smax) continue;
    const int ns = S.snx * S.sny;
    hipLaunchKernelGGL(strip_max, dim3((ns + 3) / 4), dim3(256), 0, q, S.H, S.nx, S.ny, S.snx, S.sny, S.smax);
    if (hipGetLastError() != hipSuccess) return mch::fail(MC_ERR_HIP, "strip maxima launch failed");
  }
  // sub-steps in, results out: one device buffer for every program
  size_t nin = 0, nlut = 0, nfr = 0, nwall = 0, nedge = 0, nsave = 0;
  bool flex = false;
  for (int p = 0; p < n; p++) {
    nin += (size_t)progs[p].nsub;
    if (progs[p].type == MC_TOOL_BULL || progs[p].type == MC_TOOL_PROFILE) nlut += (size_t)progs[p].lutn;
    if (progs[p].type == MC_TOOL_PROFILE) nedge += (size_t)progs[p].edge_n;
    nfr += (size_t)pd[(size_t)p].nframes;
    if (progs[p].flex && progs[p].walls) nwall += (size_t)progs[p].nsub;
    flex = flex || progs[p].flex;
    if (pd[(size_t)p].hOnly && progs[p].nsub > 0) nsave += (size_t)pd[(size_t)p].act.nx * (size_t)pd[(size_t)p].act.ny;
  }
  // the wall model's walk table, s = step, s + step, ... up to 60 mm, summed in the
  // reference's order; one per distinct step
  struct Table { double step; size_t at; int n20, n60; };
  std::vector<double> sacc;
  std::vector<Table> tables;
  std::vector<size_t> soff((size_t)n, 0);
  for (int p = 0; p < n; p++) {
    ProgDev& d = pd[(size_t)p];
    if (!d.flex) continue;
    const Table* tb = nullptr;
    for (const Table& t : tables)
      if (t.step == d.wstep) { tb = &t; break; }
    if (!tb) {
      Table t{d.wstep, sacc.size(), 0, 0};
      for (double v = d.wstep; v <= 60; v += d.wstep) {
        sacc.push_back(v);
        t.n60++;
        if (v <= 20) t.n20++;
      }
      tables.push_back(t);
      tb = &tables.back();
    }
    soff[(size_t)p] = tb->at;
    d.n20 = tb->n20; d.n60 = tb->n60;
  }
  float* dlut = nullptr;
  mc_substep* dsub = nullptr;
  mc_substep_out* dout = nullptr;
  mc_frame_out* dfr = nullptr;
  ProgDev* dprog = nullptr;
  double* dsacc = nullptr;
  mc_wall_out* dwall = nullptr;
  double* dedge = nullptr;
  double* ddyn = nullptr;
  float* dsave = nullptr;   // flat and ball programs: the actual stock's heights before the launch
  auto release = [&]() {
    if (dlut) (void)hipFree(dlut);
    if (dsub) (void)hipFree(dsub);
    if (dout) (void)hipFree(dout);
    if (dfr) (void)hipFree(dfr);
    if (dprog) (void)hipFree(dprog);
    if (dsacc) (void)hipFree(dsacc);
    if (dwall) (void)hipFree(dwall);
    if (dedge) (void)hipFree(dedge);
    if (ddyn) (void)hipFree(ddyn);
    if (dsave) (void)hipFree(dsave);
  };
  if (hipMalloc(&dsub, sizeof(mc_substep) * std::max<size_t>(1, nin)) != hipSuccess ||
      hipMalloc(&dout, sizeof(mc_substep_out) * std::max<size_t>(1, nin)) != hipSuccess ||
      hipMalloc(&dfr, sizeof(mc_frame_out) * std::max<size_t>(1, nfr)) != hipSuccess ||
      hipMalloc(&dprog, sizeof(ProgDev) * (size_t)n) != hipSuccess ||
      hipMalloc(&dlut, sizeof(float) * std::max<size_t>(1, nlut)) != hipSuccess ||
      hipMalloc(&dsacc, sizeof(double) * std::max<size_t>(1, sacc.size())) != hipSuccess ||
      hipMalloc(&dwall, sizeof(mc_wall_out) * std::max<size_t>(1, nwall)) != hipSuccess ||
      hipMalloc(&dedge, sizeof(double) * std::max<size_t>(1, nedge)) != hipSuccess ||
      hipMalloc(&ddyn, sizeof(double) * (size_t)n) != hipSuccess ||
      hipMalloc(&dsave, sizeof(float) * std::max<size_t>(1, nsave)) != hipSuccess) {
    release();
    return mch::fail(MC_ERR_NOMEM, "sub-step buffers");
  }
  size_t off = 0, loff = 0, foff = 0, woff = 0, eoff = 0;
  int st = MC_OK;
  if (!sacc.empty() &&
      hipMemcpyAsync(dsacc, sacc.data(), sizeof(double) * sacc.size(), hipMemcpyHostToDevice, q) != hipSuccess)
    st = mch::fail(MC_ERR_HIP, "wall table upload");
  for (int p = 0; p < n && !st; p++) {
    const mc_sim_program& g = progs[p];
    if (g.flex) {
      pd[(size_t)p].sacc = dsacc + soff[(size_t)p];
      if (g.walls) {
        pd[(size_t)p].walls = dwall + woff;
        woff += (size_t)g.nsub;
      }
    }
    if (g.type == MC_TOOL_PROFILE) {
      pd[(size_t)p].edge.r = dedge + eoff;
      if (hipMemcpyAsync(dedge + eoff, g.edge_r, sizeof(double) * (size_t)g.edge_n, hipMemcpyHostToDevice, q) != hipSuccess)
        st = mch::fail(MC_ERR_HIP, "edge profile upload");
      eoff += (size_t)g.edge_n;
    }
    if (g.type == MC_TOOL_BULL || g.type == MC_TOOL_PROFILE) {
      pd[(size_t)p].lut = dlut + loff;
      if (hipMemcpyAsync(dlut + loff, g.lut, sizeof(float) * (size_t)g.lutn, hipMemcpyHostToDevice, q) != hipSuccess)
        st = mch::fail(MC_ERR_HIP, "tip table upload");
      loff += (size_t)g.lutn;
    }
    pd[(size_t)p].sub = dsub + off;
    pd[(size_t)p].out = dout + off;
    if (g.budgetLeft) pd[(size_t)p].dynOut = ddyn + p;
    if (pd[(size_t)p].nframes > 0) {
      pd[(size_t)p].frames = dfr + foff;
      foff += (size_t)pd[(size_t)p].nframes;
    }
    if (g.nsub > 0 && hipMemcpyAsync(dsub + off, g.sub, sizeof(mc_substep) * (size_t)g.nsub, hipMemcpyHostToDevice, q) != hipSuccess)
      st = mch::fail(MC_ERR_HIP, "sub-step upload");
    off += (size_t)g.nsub;
    // the engine's record with its run constants, uploaded; chip events are not kept here
    if (!st) st = mce_engine_prepare(g.engine, 0);
    pd[(size_t)p].eng = mce_engine_device(g.engine);
  }
  if (!st && hipMemcpyAsync(dprog, pd.data(), sizeof(ProgDev) * (size_t)n, hipMemcpyHostToDevice, q) != hipSuccess)
    st = mch::fail(MC_ERR_HIP, "program upload");
  std::vector<size_t> saveAt((size_t)n, 0);
  for (size_t p = 0, at = 0; p < (size_t)n && !st; p++) {
    const StockDev& A = pd[p].act;
    if (!pd[p].hOnly || progs[p].nsub == 0) continue;
    saveAt[p] = at;
    if (hipMemcpyAsync(dsave + at, A.H, sizeof(float) * (size_t)A.nx * A.ny, hipMemcpyDeviceToDevice, q) != hipSuccess)
      st = mch::fail(MC_ERR_HIP, "height save");
    at += (size_t)A.nx * A.ny;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = mch::fail(MC_ERR_HIP, "timing events");
  if (!st) {
    (void)hipEventRecord(e0, q);
    if (flex)
      hipLaunchKernelGGL(sim_run<true>, dim3(n), dim3(NT), 0, q, (const ProgDev*)dprog);
    else
      hipLaunchKernelGGL(sim_run<false>, dim3(n), dim3(NT), 0, q, (const ProgDev*)dprog);
    if (hipGetLastError() != hipSuccess) st = mch::fail(MC_ERR_HIP, "sub-step pipeline launch failed");
    (void)hipEventRecord(e1, q);
  }
  // once the kernel has run, every engine's host mirror is refreshed from its device record,
  // even after a failed download, so that no mirror keeps the state from before the launch
  const bool launched = !st;
  off = 0;
  foff = 0;
  woff = 0;
  std::vector<mc_substep_out> hout(std::max<size_t>(1, nin));
  for (int p = 0; p < n && launched; p++) {
    const mc_sim_program& g = progs[p];
    if (!st && (g.out || pd[(size_t)p].hOnly) && g.nsub > 0 &&
        hipMemcpyAsync(hout.data() + off, dout + off, sizeof(mc_substep_out) * (size_t)g.nsub, hipMemcpyDeviceToHost,
                       q) != hipSuccess)
      st = mch::fail(MC_ERR_HIP, "result download");
    off += (size_t)g.nsub;
    const int nf = pd[(size_t)p].nframes;
    if (!st && nf > 0 &&
        hipMemcpyAsync(g.frames, dfr + foff, sizeof(mc_frame_out) * (size_t)nf, hipMemcpyDeviceToHost, q) != hipSuccess)
      st = mch::fail(MC_ERR_HIP, "frame download");
    foff += (size_t)nf;
    if (g.flex && g.walls) {
      if (!st && g.nsub > 0 &&
          hipMemcpyAsync(g.walls, dwall + woff, sizeof(mc_wall_out) * (size_t)g.nsub, hipMemcpyDeviceToHost, q) != hipSuccess)
        st = mch::fail(MC_ERR_HIP, "wall record download");
      woff += (size_t)g.nsub;
    }
    if (!st && g.budgetLeft &&
        hipMemcpyAsync(g.budgetLeft, ddyn + p, sizeof(double), hipMemcpyDeviceToHost, q) != hipSuccess)
      st = mch::fail(MC_ERR_HIP, "budget download");
    const int fst = mce_engine_finish(g.engine);
    if (!st) st = fst;
  }
  if (!st && hipStreamSynchronize(q) != hipSuccess) st = mch::fail(MC_ERR_HIP, "sub-step pipeline failed");
  float ms = 0;
  if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 0;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  // Flat and ball programs: the chain wrote the actual stock's heights alone.  Its deflected stamps
  // now run as one batch with Stock.stamp's sequential semantics (src/stock.js:118-164) from
  // the heights saved before the launch, writing the heights again (the same values) with the
  // errors, wall distances, lines, feed marks and dirty box, and giving each stamp's volume.
  // The stamp is the chain's: at p plus the deflection, R less the radial wear, the sub-step's
  // error value and line, the feed mark of the last cutting sub-step.
  lk.unlock();
  off = 0;
  for (int p = 0; p < n && !st; p++) {
    const mc_sim_program& g = progs[p];
    if (pd[(size_t)p].hOnly && g.nsub > 0) {
      const StockDev& A = pd[(size_t)p].act;
      std::vector<mc_stamp> cmd((size_t)g.nsub);
      std::vector<mc_stamp_opt> opt((size_t)g.nsub);
      std::vector<double> vol((size_t)g.nsub);
      double fzl = g.fz;
      for (int k = 0; k < g.nsub; k++) {
        const mc_substep& s = g.sub[k];
        const mc_substep_out& o = hout[off + (size_t)k];
        if (o.cut) fzl = s.fzEff;
        const bool bl = g.type == MC_TOOL_BALL;
        cmd[(size_t)k] = mc_stamp{s.x + o.dx, s.y + o.dy, s.z + (bl ? g.rw : 0.0), g.R - g.rw,
                                  bl ? MC_TOOL_BALL : MC_TOOL_FLAT, s.line, o.errVal};
        opt[(size_t)k] = mc_stamp_opt{s.ux, s.uy, fzl * pd[(size_t)p].N, -1, 0};
      }
      {
        std::lock_guard<std::mutex> lk2(mch::lock());
        if (hipMemcpyAsync(A.H, dsave + saveAt[(size_t)p], sizeof(float) * (size_t)A.nx * A.ny, hipMemcpyDeviceToDevice,
                           q) != hipSuccess ||
            hipStreamSynchronize(q) != hipSuccess)
          st = mch::fail(MC_ERR_HIP, "height restore");
      }
      if (!st) st = mc_stamp_batch_ex(g.actual, cmd.data(), opt.data(), g.nsub, nullptr, nullptr, nullptr, 0, vol.data());
      for (int k = 0; k < g.nsub && !st; k++) hout[off + (size_t)k].vol = vol[(size_t)k];
    }
    if (!st && g.out && g.nsub > 0)
      std::memcpy(g.out, hout.data() + off, sizeof(mc_substep_out) * (size_t)g.nsub);
    off += (size_t)g.nsub;
  }
  if (!st) mch::set_last_ms(ms);
  {
    std::lock_guard<std::mutex> lk2(mch::lock());
    release();
  }
  return st;
}
