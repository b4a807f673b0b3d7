// stock.hip — heightmap stock removal (Stock.stamp, src/stock.js:118-164) and the
// stock's queries (blockMax, maxInDisk, sample: src/stock.js:45-92) on gfx950.
//
// The heightfield and its side maps stay resident in HBM.  A batch of n stamps is applied
// with the reference's sequential semantics by giving every grid node one thread that
// visits, in stamp order, exactly the stamps whose bounding box covers it; nodes are
// independent, so the final H, err, wallD, lineCut, lineWall, markS and markF are
// bit-identical to n successive Stock.stamp calls, while each node's state is read from
// and written to HBM once per batch (the tile keeps it in registers).
//
//   stamp_prep   per stamp: bounding box (the reference's floor/ceil rule), radii, rim
//                height and lowest reach; per chunk of CHUNK stamps: union bounding box
//                and lowest reach.
//   stamp_tiles  one 16x16 node tile per workgroup: lists, 256 chunks at a time, the chunks
//                whose box meets the tile and whose lowest reach is below the tile's
//                highest node, walks them, compacts (in order, with wave ballots) the stamps of a chunk that
//                meet the tile and can still change it into LDS, and lets each node thread
//                apply them in order.  A stamp can change a node only where H > zmin (its
//                lowest reach: the cut, the wall test and the feed-mark test all need it),
//                so stamps whose zmin is at or above the tile's highest node are dropped,
//                and each wave skips the stamps at or above its own highest node: exact,
//                since those stamps would change nothing.  The EX variant adds tip
//                profiles and feed marks.
//
// Stock.dirty (src/stock.js:156-160), the union of the boxes of the stamps that changed a
// node, is kept per wavefront while it walks the stamps and folded into the stock's rectangle
// by four integer atomics per wavefront at the end (min / max: exact in any order).
//
// Removed volume: each wavefront's share of a stamp goes into an LDS partial of the tile
// (one slot per compacted stamp of the chunk), and the tile adds its partials to the output
// once per chunk: per stamp (the value stamp() returns), or per program line (the
// simulator's per-segment removed volume; each wave folds its 64 slots of the chunk line by
// line, so the stamps of one line cost one global atomic per wave and chunk).  The FP64 sums' last bits depend on the summation order.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mcsim.h"
#include "mc_math.h"

namespace mcs {

constexpr int TILE = 16;      // 16 x 16 nodes per workgroup
constexpr int CHUNK = 256;    // stamps compacted into LDS at a time
constexpr double EDGE_TIE = 2 * 0.0015;  // src/stock.js:6
constexpr double LUT_DR = MC_LUT_DR;     // src/assembly.js:8

struct StampDev {
  double cx, cy, tipZ, R, R2, tie, Rw2, zRim;
  double zmin;   // lowest z the stamp can reach: every update it makes needs H above it
  float errVal;
  int32_t line, type, i0, i1, j0, j1, index;
};

// tip profile and feed marks of a stamp (the EX variant only)
struct StampX {
  double ux, uy, mf;       // mf: mark spacing, 0 = no marks (src/stock.js:128)
  int32_t lut_off, lut_n;  // profile table in the batch's LUT buffer, lut_n = 0: none
};

struct ChunkBox {
  int32_t i0, i1, j0, j1;
  float zmin;        // lowest reach of the chunk's stamps, rounded down to f32
  int32_t pad[3];
};

__global__ void stamp_prep(const mc_stamp* __restrict__ in, const mc_stamp_opt* __restrict__ opt,
                           const float* __restrict__ luts, const int32_t* __restrict__ lut_off,
                           const int32_t* __restrict__ lut_len, const float* __restrict__ lut_min,
                           int n, double hx, double hy, int nx, int ny, double zTop, double band,
                           StampDev* __restrict__ out, StampX* __restrict__ outx,
                           int4* __restrict__ cull, float* __restrict__ cull_z, ChunkBox* __restrict__ box) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ int bi0[256], bi1[256], bj0[256], bj1[256];
  __shared__ float bz[256];
  int i0 = 1 << 30, i1 = -1, j0 = 1 << 30, j1 = -1;
  float zlow = __builtin_inff();
  if (s < n) {
    const mc_stamp st = in[s];
    StampDev d;
    StampX x{0, 0, 0, 0, 0};
    const int t = opt ? opt[s].lut : -1;
    if (t >= 0) { x.lut_off = lut_off[t]; x.lut_n = lut_len[t]; }
    d.cx = st.cx; d.cy = st.cy; d.tipZ = st.tipZ; d.R = st.R;
    d.type = st.type; d.line = st.line; d.errVal = (float)st.errVal; d.index = s;
    const double m = 2 * (hx > hy ? hx : hy);
    const double Rw = st.R + (m > band ? m : band);
    d.Rw2 = Rw * Rw;
    d.R2 = st.R * st.R;
    d.tie = d.R2 - EDGE_TIE * st.R;
    const bool ball = st.type == MC_TOOL_BALL;
    d.zRim = st.tipZ + (t >= 0 ? (double)luts[x.lut_off + x.lut_n - 1] : ball ? st.R : 0);
    if (opt && opt[s].f > 0 && st.type == MC_TOOL_FLAT && t < 0) {
      x.ux = opt[s].ux; x.uy = opt[s].uy; x.mf = opt[s].f;
    }
    // lowest reach: a profile cuts to tipZ + lut[q] >= tipZ + min(lut) (rounding is
    // monotone), a ball to (tipZ + R) - sqrt(R^2 - r2) >= (tipZ + R) - R, a flat one to
    // tipZ, and a flat one with feed marks rewrites the marks where H > tipZ - 0.002; the
    // wall update needs H > zRim >= zmin too, so a node with H <= zmin is left as it is
    d.zmin = t >= 0 ? st.tipZ + (double)lut_min[t]
           : ball ? (st.tipZ + st.R) - st.R
           : x.mf != 0 ? st.tipZ - 0.002 : st.tipZ;
    const double a = __builtin_floor((st.cx - Rw) / hx), b = __builtin_ceil((st.cx + Rw) / hx);
    const double c = __builtin_floor((st.cy - Rw) / hy), e = __builtin_ceil((st.cy + Rw) / hy);
    const bool live = st.tipZ < zTop && a <= nx - 1 && c <= ny - 1 && b >= 0 && e >= 0;
    d.i0 = live ? (int)(a > 0 ? a : 0) : 1;
    d.i1 = live ? (int)(b < nx - 1 ? b : nx - 1) : 0;
    d.j0 = live ? (int)(c > 0 ? c : 0) : 1;
    d.j1 = live ? (int)(e < ny - 1 ? e : ny - 1) : 0;
    out[s] = d;
    // compact copy for the tiles' culling pass: the box, and zmin rounded down to f32 (for
    // an f32 height h, h > zmin exactly when h > RD_f32(zmin))
    cull[s] = make_int4(d.i0, d.i1, d.j0, d.j1);
    cull_z[s] = __double2float_rd(d.zmin);
    zlow = cull_z[s];
    if (outx) outx[s] = x;
    if (d.i0 <= d.i1 && d.j0 <= d.j1) { i0 = d.i0; i1 = d.i1; j0 = d.j0; j1 = d.j1; }
  }
  bi0[threadIdx.x] = i0; bi1[threadIdx.x] = i1; bj0[threadIdx.x] = j0; bj1[threadIdx.x] = j1;
  bz[threadIdx.x] = zlow;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int t = threadIdx.x;
      bi0[t] = min(bi0[t], bi0[t + w]); bi1[t] = max(bi1[t], bi1[t + w]);
      bj0[t] = min(bj0[t], bj0[t + w]); bj1[t] = max(bj1[t], bj1[t + w]);
      bz[t] = fminf(bz[t], bz[t + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) box[blockIdx.x] = ChunkBox{bi0[0], bi1[0], bj0[0], bj1[0], bz[0], {0, 0, 0}};
}

// EX: tip profiles (luts) and feed marks (markS / markF) are live in this batch; LINES: vol
// is indexed by the stamps' program line instead of their index
template <bool EX, bool LINES>
__global__ __launch_bounds__(TILE * TILE) void stamp_tiles(
    const StampDev* __restrict__ st, const StampX* __restrict__ stx, const int4* __restrict__ cull,
    const float* __restrict__ cull_z, const ChunkBox* __restrict__ box, int n, int nx, int ny, double hx, double hy, double zB, float* __restrict__ H,
    float* __restrict__ err, float* __restrict__ wallD, int32_t* __restrict__ lineCut,
    int32_t* __restrict__ lineWall, float* __restrict__ markS, float* __restrict__ markF,
    const float* __restrict__ luts, int32_t* __restrict__ rect, double* __restrict__ vol) {
  __shared__ StampDev list[CHUNK];
  __shared__ double cvol[CHUNK];   // the tile's removed height per compacted stamp of the chunk
  __shared__ StampX listx[EX ? CHUNK : 1];
  const int t = threadIdx.x;
  const int ti0 = blockIdx.x * TILE, tj0 = blockIdx.y * TILE;
  const int ti1 = min(ti0 + TILE, nx) - 1, tj1 = min(tj0 + TILE, ny) - 1;
  const int i = ti0 + (t % TILE), j = tj0 + (t / TILE);
  const bool own = i <= ti1 && j <= tj1;
  const int k = j * nx + i;
  float h = 0, e = 0, wd = 0, ms = 0, mfv = 0;
  int lc = 0, lw = 0;
  int ri0 = 0x7fffffff, rj0 = 0x7fffffff, ri1 = -1, rj1 = -1;   // boxes of the stamps this wave changed
  if (own) {
    h = H[k]; e = err[k]; wd = wallD[k]; lc = lineCut[k]; lw = lineWall[k];
    if (EX) { ms = markS[k]; mfv = markF[k]; }
  }
  const double xi = i * hx, yj = j * hy;
  const int nchunks = (n + CHUNK - 1) / CHUNK;
  const int lane = t & 63, wv = t >> 6;
  __shared__ int wcount[TILE * TILE / 64];
  __shared__ float wmaxs[TILE * TILE / 64];
  __shared__ int clist[TILE * TILE];   // chunks that meet the tile, in order
  __shared__ float clz[TILE * TILE];   // and their lowest reach
  __shared__ int ccount[TILE * TILE / 64];
  for (int c0 = 0; c0 < nchunks; c0 += TILE * TILE) {
    // the boxes of the next 256 chunks are read at once (one per thread) and the chunks that
    // meet the tile listed in order, so that walking the chunk list costs one load latency
    // per 256 chunks instead of one per chunk
    {
      // the tile's highest node now (H only falls): a chunk whose lowest reach is at or
      // above it can change no node of the tile, now or later
      float wm = own ? h : -__builtin_inff();
      for (int o = 32; o > 0; o >>= 1) wm = fmaxf(wm, __shfl_xor(wm, o, 64));
      if (lane == 0) wmaxs[wv] = wm;
      const int cc = c0 + t;
      bool meets = false;
      ChunkBox b{};
      if (cc < nchunks) {
        b = box[cc];
        meets = !(b.i1 < ti0 || b.i0 > ti1 || b.j1 < tj0 || b.j0 > tj1);
      }
      __syncthreads();
      float tm = wmaxs[0];
      for (int v = 1; v < TILE * TILE / 64; v++) tm = fmaxf(tm, wmaxs[v]);
      meets = meets && tm > b.zmin;
      const unsigned long long cm = __builtin_amdgcn_ballot_w64(meets);
      if (lane == 0) ccount[wv] = __builtin_popcountll(cm);
      __syncthreads();
      int cbefore = 0;
      for (int v = 0; v < wv; v++) cbefore += ccount[v];
      if (meets) {
        const int at = cbefore + __builtin_amdgcn_mbcnt_hi((unsigned)(cm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)cm, 0));
        clist[at] = cc;
        clz[at] = b.zmin;
      }
      __syncthreads();
    }
    int ctotal = 0;
    for (int v = 0; v < TILE * TILE / 64; v++) ctotal += ccount[v];
    for (int k = 0; k < ctotal; k++) {
      const int c = clist[k];
      // highest node of this wave and of the tile (H only falls, so a stale maximum is a
      // safe upper bound for the rest of the chunk)
      float wmax = own ? h : -__builtin_inff();
      for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
      if (lane == 0) wmaxs[wv] = wmax;
      // order-preserving compaction of the chunk's stamps that meet this tile and reach
      // below its highest node
      const int s = c * CHUNK + t;
      __syncthreads();
      float tmax = wmaxs[0];
      for (int v = 1; v < TILE * TILE / 64; v++) tmax = fmaxf(tmax, wmaxs[v]);
      // the tile may have fallen below the chunk's lowest reach since the list was made
      if (!(tmax > clz[k])) { __syncthreads(); continue; }   // block-uniform
      bool cand = false;
      float zmin = 0;
      if (s < n) {
        const int4 b4 = cull[s];
        cand = !(b4.y < ti0 || b4.x > ti1 || b4.w < tj0 || b4.z > tj1) && b4.x <= b4.y && b4.z <= b4.w;
        zmin = cull_z[s];
      }
      const bool hit = cand && tmax > zmin;
      const unsigned long long m = __builtin_amdgcn_ballot_w64(hit);
      if (lane == 0) wcount[wv] = __builtin_popcountll(m);
      __syncthreads();
      int before = 0, total = 0;
      for (int v = 0; v < TILE * TILE / 64; v++) {
        before += v < wv ? wcount[v] : 0;
        total += wcount[v];
      }
      if (hit) {
        const int slot = before + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
        list[slot] = st[s];
        if (EX) listx[slot] = stx[s];
      }
      cvol[t] = 0.0;
      __syncthreads();
      for (int q = 0; q < total; q++) {
        const StampDev& p = list[q];
        if ((double)wmax <= p.zmin) continue;   // wave-uniform: no node of this wave can change
        double removed = 0;
        bool changed = false;
        if (own && i >= p.i0 && i <= p.i1 && j >= p.j0 && j <= p.j1) {
          const double dx = xi - p.cx, dy = yj - p.cy;
          const double r2 = dx * dx + dy * dy;
          if (r2 >= p.tie) {
            if (r2 <= p.Rw2 && (double)h > p.zRim) {
              const double dd = __builtin_sqrt(r2) - p.R;
              if (dd < (double)wd) { wd = (float)dd; lw = p.line; changed = true; }
            }
          } else {
            double z;
            if (EX && listx[q].lut_n > 0) {
              const StampX& x = listx[q];
              const double r = mcm::round_js(__builtin_sqrt(r2) / LUT_DR);
              z = p.tipZ + (double)luts[x.lut_off + (r < x.lut_n - 1 ? (int)r : x.lut_n - 1)];
            } else {
              z = p.type == MC_TOOL_BALL ? p.tipZ + p.R - __builtin_sqrt(p.R2 - r2) : p.tipZ;
            }
            if ((double)h > z) {
              removed = (double)h - z;
              h = (float)(z > zB ? z : zB);
              e = p.errVal;
              wd = __builtin_inff();
              lc = p.line;
              changed = true;
            }
            // the last pass over a floor at this depth leaves its trailing-edge arcs
            if (EX && listx[q].mf != 0 && (double)h > z - 0.002) {
              const StampX& x = listx[q];
              const double pp = -dx * x.uy + dy * x.ux;
              const double u = p.R2 - pp * pp;
              ms = (float)(xi * x.ux + yj * x.uy + __builtin_sqrt(u > 0 ? u : 0.0));
              mfv = (float)x.mf;
              changed = true;
            }
          }
        }
        if (__any(changed)) {   // wave-uniform: the stamp's box joins the dirty rectangle
          ri0 = min(ri0, p.i0); rj0 = min(rj0, p.j0); ri1 = max(ri1, p.i1); rj1 = max(rj1, p.j1);
        }
        if (vol && __any(removed != 0.0)) {
          double r = removed;
          for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
          if (lane == 0 && r != 0.0) atomicAdd(&cvol[q], r);
        }
      }
      __syncthreads();
      // the tile's partial volumes of this chunk, added to the output once
      if (vol) {
        if (!LINES) {
          if (t < total && cvol[t] != 0.0) atomicAdd(&vol[list[t].index], cvol[t]);
        } else {
          // each wave folds its 64 slots line by line: one atomic per distinct line (the
          // stamps of a chunk mostly share one or two)
          bool live = t < total && cvol[t] != 0.0;
          const int L = live ? list[t].line : -1;
          const double v = live ? cvol[t] : 0.0;
          for (unsigned long long m = __builtin_amdgcn_ballot_w64(live); m != 0;
               m = __builtin_amdgcn_ballot_w64(live)) {   // wave-uniform
            const int first = __builtin_ctzll(m);
            const int Lf = __shfl(L, first, 64);
            const bool mine = live && L == Lf;
            double r = mine ? v : 0.0;
            for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
            if (lane == first) atomicAdd(&vol[Lf], r);
            live = live && !mine;
          }
        }
      }
      __syncthreads();
    }
    __syncthreads();   // every wave has read ccount and clist before the next 256 chunks
  }
  if (own) {
    H[k] = h; err[k] = e; wallD[k] = wd; lineCut[k] = lc; lineWall[k] = lw;
    if (EX) { markS[k] = ms; markF[k] = mfv; }
  }
  if (lane == 0 && ri1 >= 0) {
    atomicMin(&rect[0], ri0); atomicMin(&rect[1], rj0);
    atomicMax(&rect[2], ri1); atomicMax(&rect[3], rj1);
  }
}

__global__ void set_rect(int32_t* rect) {
  if (threadIdx.x == 0) { rect[0] = 0x7fffffff; rect[1] = 0x7fffffff; rect[2] = -1; rect[3] = -1; }
}

}  // namespace mcs

namespace mcs {

// Stock.sample (src/stock.js:85-92): bilinear height, -inf outside the stock
__device__ __forceinline__ double sample_h(const float* __restrict__ H, int nx, int ny, double hx,
                                          double hy, double x, double y) {
  const double fx = x / hx, fy = y / hy;
  if (fx < 0 || fy < 0 || fx > nx - 1 || fy > ny - 1) return -__builtin_inf();
  int i = (int)fx, j = (int)fy;
  if (i > nx - 2) i = nx - 2;
  if (j > ny - 2) j = ny - 2;
  const double tx = fx - i, ty = fy - j;
  const double a = H[j * nx + i], b = H[j * nx + i + 1], c = H[(j + 1) * nx + i],
               d = H[(j + 1) * nx + i + 1];
  return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}

__global__ void stock_sample(const float* __restrict__ H, int nx, int ny, double hx, double hy,
                             const double* __restrict__ xy, int n, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  out[q] = sample_h(H, nx, ny, hx, hy, xy[2 * q], xy[2 * q + 1]);
}

// Simulator.fillEngagement (src/sim.js:301-319), immersion (:343-348) and the axial depth
// (:387-388) for one query per wavefront: lane b (and b + 64, b + 128) samples the stock on
// the cutter periphery for angle bin b.  Hb is stored as f32 like engine.Hb; the maximum
// and the first / last leading bin above tipZ + 0.1 are wave reductions, exact since max
// and min of f32 values and of bin indices do not depend on order.
__global__ __launch_bounds__(256) void engagement(
    const float* __restrict__ H, int nx, int ny, double hx, double hy,
    const mc_engage_query* __restrict__ qs, int nq, float* __restrict__ Hb_out,
    mc_engage_result* __restrict__ out) {
  constexpr double PI = 3.141592653589793;
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (w >= nq) return;   // wave-uniform
  const mc_engage_query q = qs[w];
  const double cell = hx > hy ? hx : hy;
  const double r = q.R - 0.5 * cell, ax = q.cx + q.ux * cell, ay = q.cy + q.uy * cell;
  const double vx = -q.uy, vy = q.ux;
  const double z = q.tipZ + 0.1;
  float hmax = -__builtin_inff();
  int first = MC_BINS, last = -1;
  for (int b = lane; b < MC_BINS; b += 64) {
    const double phi = (b + 0.5) / MC_BINS * 2 * PI;
    double sn, cs;
    mcm::sincos_js(phi, sn, cs);
    const double h = sample_h(H, nx, ny, hx, hy, ax + r * (sn * q.ux + cs * vx), ay + r * (sn * q.uy + cs * vy));
    const float f = sn > 0.02 ? (float)h : -__builtin_inff();
    if (Hb_out) Hb_out[(size_t)w * MC_BINS + b] = f;
    hmax = fmaxf(hmax, f);
    if (b < MC_BINS / 2 && (double)f > z) { first = min(first, b); last = max(last, b); }
  }
  for (int o = 32; o > 0; o >>= 1) {
    hmax = fmaxf(hmax, __shfl_xor(hmax, o, 64));
    first = min(first, __shfl_xor(first, o, 64));
    last = max(last, __shfl_xor(last, o, 64));
  }
  if (lane == 0) {
    mc_engage_result res;
    res.hmax = hmax;
    const double d = (double)hmax - q.tipZ;
    const double a = d > 0 ? d : 0;     // max(0, -inf) = 0
    res.ap = q.fluteLen < a ? q.fluteLen : a;
    res.imm = last >= 0;
    res.pad = 0;
    res.phiSt = 0;
    res.phiEx = 0;
    if (last >= 0) {
      const double st = (first + 0.5) / MC_BINS * 2 * PI - PI / MC_BINS;
      const double ex = (last + 0.5) / MC_BINS * 2 * PI + PI / MC_BINS;
      res.phiSt = st > 0 ? st : 0;
      res.phiEx = ex < PI ? ex : PI;
    }
    out[w] = res;
  }
}

constexpr int BLK = MC_STOCK_B;   // block edge of the interference maxima (src/stock.js:35)

// Stock.blockMax (src/stock.js:45-54) for every block: the reference's scan order and
// strict '>' (so which of +0 / -0 wins is the reference's too)
__global__ void block_max(const float* __restrict__ H, int nx, int ny, int bnx, int bny,
                          float* __restrict__ bmax) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bnx * bny) return;
  const int bi = b % bnx, bj = b / bnx;
  float m = -__builtin_inff();
  const int ie = min(nx, bi * BLK + BLK), je = min(ny, bj * BLK + BLK);
  for (int j = bj * BLK; j < je; j++)
    for (int i = bi * BLK; i < ie; i++) {
      const float v = H[j * nx + i];
      if (v > m) m = v;
    }
  bmax[b] = m;
}

// Stock.maxInDisk (src/stock.js:57-80), one wavefront per query.  The reference walks the
// blocks of the disk's bounding box in row-major order and lets a block replace the best
// answer when its maximum m exceeds max(best, above): a block wholly inside the disk
// answers m at its centre, an edge block its highest node inside the disk (first in scan
// order).  Since m >= every node of its block, this equals: among the blocks with
// m > above, the first (in walk order) whose candidate is the largest, provided that
// candidate exceeds -inf.  Lanes take 64 blocks at a time; a (value, order) reduction
// with strict '>' keeps the first of equal candidates.
__global__ __launch_bounds__(256) void max_in_disk(
    const float* __restrict__ H, const float* __restrict__ bmax, int nx, int ny, int bnx,
    double hx, double hy, const double* __restrict__ q, int nq, double* __restrict__ out) {
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (w >= nq) return;   // wave-uniform
  const double cx = q[4 * w], cy = q[4 * w + 1], r = q[4 * w + 2], above = q[4 * w + 3];
  double a = __builtin_floor((cx - r) / hx), b = __builtin_ceil((cx + r) / hx);
  double c = __builtin_floor((cy - r) / hy), d = __builtin_ceil((cy + r) / hy);
  a = a > 0 ? a : 0; b = b < nx - 1 ? b : nx - 1;
  c = c > 0 ? c : 0; d = d < ny - 1 ? d : ny - 1;
  double best = -__builtin_inf(), bx = 0, by = 0;
  if (a <= b && c <= d) {
    const int i0 = (int)a, i1 = (int)b, j0 = (int)c, j1 = (int)d;
    const int bi0 = i0 / BLK, bj0 = j0 / BLK;
    const int nbi = i1 / BLK - bi0 + 1, nb = nbi * (j1 / BLK - bj0 + 1);
    const double r2 = r * r;
    for (int base = 0; base < nb; base += 64) {
      const int t = base + lane;
      double v = -__builtin_inf(), vx = 0, vy = 0;
      if (t < nb) {
        const int bi = bi0 + t % nbi, bj = bj0 + t / nbi;
        const double m = bmax[bj * bnx + bi];
        if (m > above) {
          const double xa = (double)(bi * BLK) * hx - cx, xb = (double)min(nx - 1, bi * BLK + BLK - 1) * hx - cx;
          const double ya = (double)(bj * BLK) * hy - cy, yb = (double)min(ny - 1, bj * BLK + BLK - 1) * hy - cy;
          const double far = (xa * xa > xb * xb ? xa * xa : xb * xb) + (ya * ya > yb * yb ? ya * ya : yb * yb);
          if (far <= r2) {
            v = m; vx = cx + (xa + xb) / 2; vy = cy + (ya + yb) / 2;
          } else {
            const int ja = max(j0, bj * BLK), jb = min(j1, bj * BLK + BLK - 1);
            const int ia = max(i0, bi * BLK), ib = min(i1, bi * BLK + BLK - 1);
            for (int j = ja; j <= jb; j++) {
              const double dy = j * hy - cy;
              for (int i = ia; i <= ib; i++) {
                const double dx = i * hx - cx, h = H[j * nx + i];
                if (h > v && dx * dx + dy * dy <= r2) { v = h; vx = i * hx; vy = j * hy; }
              }
            }
          }
        }
      }
      // first of the largest: lane order is walk order
      int order = t;
      for (int o = 1; o < 64; o <<= 1) {
        const double ov = __shfl_xor(v, o, 64), ox = __shfl_xor(vx, o, 64), oy = __shfl_xor(vy, o, 64);
        const int oo = __shfl_xor(order, o, 64);
        if (ov > v || (ov == v && oo < order)) { v = ov; vx = ox; vy = oy; order = oo; }
      }
      if (v > best) { best = v; bx = vx; by = vy; }
    }
  }
  if (lane == 0) { out[3 * w] = best; out[3 * w + 1] = bx; out[3 * w + 2] = by; }
}

__global__ void fill_f32(float* p, size_t n, float v) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = v;
}

}  // namespace mcs

// ---- scratch of the no-allocation stamp batch (mc_stamp_batch_noalloc)
// The uploaded stamps, their expanded and culling records, the chunk boxes and the per-stamp
// volumes of at most MC_NOALLOC_MAX_STAMPS stamps, in one file-scope
// device array.  Each device the library runs on holds its own copy until the process ends;
// on the host the array's name is only a stand-in, so kernels get the address that
// hipGetSymbolAddress returns for the current device (resolved once per device at set-up).
namespace mcs {
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_MAX = MC_NOALLOC_MAX_STAMPS;
constexpr size_t NA_CHUNKS = (NA_MAX + CHUNK - 1) / CHUNK;
constexpr size_t NA_OFF_DEV = align256(sizeof(mc_stamp) * NA_MAX);
constexpr size_t NA_OFF_CULL = align256(NA_OFF_DEV + sizeof(StampDev) * NA_MAX);
constexpr size_t NA_OFF_BOX = align256(NA_OFF_CULL + (sizeof(int4) + sizeof(float)) * NA_MAX);
constexpr size_t NA_OFF_VOL = align256(NA_OFF_BOX + sizeof(ChunkBox) * NA_CHUNKS);
constexpr size_t NA_BYTES = align256(NA_OFF_VOL + sizeof(double) * NA_MAX);
constexpr int NA_DEVICES = 64;
static_assert(NA_BYTES <= ((size_t)MC_NOALLOC_SCRATCH_MB << 20), "scratch larger than the documented size");
}  // namespace mcs

static __device__ __align__(256) unsigned char scratch[mcs::NA_BYTES];

namespace {
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[mcs::NA_DEVICES] = {};
}  // namespace

// resolve this device's copy of the scratch array; called with `device` current
int mcs_scratch_setup(int device) {
  if (device < 0 || device >= mcs::NA_DEVICES) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -2;
  g_scratch_ptr[device] = (unsigned char*)p;
  return 0;
}

// the regions of the device's scratch (null when set-up has not resolved it)
int mcs_scratch_regions(int device, mc_stamp** in, void** stampdev, void** cull, void** box, double** vol) {
  if (device < 0 || device >= mcs::NA_DEVICES) return -1;
  unsigned char* p;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch_ptr[device];
  }
  if (!p) return -1;
  *in = (mc_stamp*)p;
  *stampdev = p + mcs::NA_OFF_DEV;
  *cull = p + mcs::NA_OFF_CULL;
  *box = p + mcs::NA_OFF_BOX;
  *vol = (double*)(p + mcs::NA_OFF_VOL);
  return 0;
}

size_t mcs_scratch_bytes() { return mcs::NA_BYTES; }

// ---- host-side launchers used by api.cpp
size_t mcs_stampdev_bytes() { return sizeof(mcs::StampDev); }
size_t mcs_chunkbox_bytes() { return sizeof(mcs::ChunkBox); }
size_t mcs_stampx_bytes() { return sizeof(mcs::StampX); }
size_t mcs_cull_bytes() { return sizeof(int4) + sizeof(float); }
int mcs_chunk() { return mcs::CHUNK; }

int mcs_launch_stamps(const mc_stamp* d_in, const mc_stamp_opt* d_opt, int n, const float* luts,
                      const int32_t* lut_off, const int32_t* lut_len, const float* lut_min,
                      void* d_stampdev, void* d_stampx, void* d_cull, void* d_box, int nx, int ny,
                      double hx, double hy, double zTop, double zB, double band, float* H,
                      float* err, float* wallD, int32_t* lineCut, int32_t* lineWall, float* markS,
                      float* markF, int32_t* rect, double* vol, bool vol_lines, hipStream_t stream, hipEvent_t ev0,
                      hipEvent_t ev1) {
  if (n <= 0) return 0;
  const int chunks = (n + mcs::CHUNK - 1) / mcs::CHUNK;
  const bool ex = d_opt != nullptr;
  hipLaunchKernelGGL(mcs::stamp_prep, dim3(chunks), dim3(mcs::CHUNK), 0, stream, d_in, d_opt, luts,
                     lut_off, lut_len, lut_min, n, hx, hy, nx, ny, zTop, band,
                     (mcs::StampDev*)d_stampdev, ex ? (mcs::StampX*)d_stampx : nullptr, (int4*)d_cull,
                     (float*)((int4*)d_cull + n), (mcs::ChunkBox*)d_box);
  const dim3 grid((nx + mcs::TILE - 1) / mcs::TILE, (ny + mcs::TILE - 1) / mcs::TILE);
  if (ev0 && hipEventRecord(ev0, stream) != hipSuccess) return -2;
  auto kern = ex ? (vol_lines ? mcs::stamp_tiles<true, true> : mcs::stamp_tiles<true, false>)
                 : (vol_lines ? mcs::stamp_tiles<false, true> : mcs::stamp_tiles<false, false>);
  hipLaunchKernelGGL(kern, grid, dim3(mcs::TILE * mcs::TILE), 0, stream,
                     (const mcs::StampDev*)d_stampdev, ex ? (const mcs::StampX*)d_stampx : (const mcs::StampX*)nullptr,
                     (const int4*)d_cull, (const float*)((int4*)d_cull + n), (const mcs::ChunkBox*)d_box, n, nx, ny,
                     hx, hy, zB, H, err, wallD, lineCut, lineWall, markS, markF, luts, rect, vol);
  if (ev1 && hipEventRecord(ev1, stream) != hipSuccess) return -2;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_set_rect(int32_t* rect, hipStream_t stream) {
  hipLaunchKernelGGL(mcs::set_rect, dim3(1), dim3(64), 0, stream, rect);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_block_max(const float* H, int nx, int ny, int bnx, int bny, float* bmax,
                         hipStream_t stream) {
  const int nb = bnx * bny;
  hipLaunchKernelGGL(mcs::block_max, dim3((nb + 255) / 256), dim3(256), 0, stream, H, nx, ny, bnx,
                     bny, bmax);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_max_in_disk(const float* H, const float* bmax, int nx, int ny, int bnx, double hx,
                           double hy, const double* d_q, int n, double* d_out, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(mcs::max_in_disk, dim3((n + 3) / 4), dim3(256), 0, stream, H, bmax, nx, ny,
                     bnx, hx, hy, d_q, n, d_out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_sample(const float* H, int nx, int ny, double hx, double hy, const double* d_xy,
                      int n, double* d_out, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(mcs::stock_sample, dim3((n + 255) / 256), dim3(256), 0, stream, H, nx, ny,
                     hx, hy, d_xy, n, d_out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_engagement(const float* H, int nx, int ny, double hx, double hy,
                          const mc_engage_query* d_q, int n, float* d_Hb, mc_engage_result* d_out,
                          hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(mcs::engagement, dim3((n + 3) / 4), dim3(256), 0, stream, H, nx, ny, hx, hy, d_q,
                     n, d_Hb, d_out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_fill(float* p, size_t n, float v, hipStream_t stream) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(mcs::fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
