// stock.hip — heightmap stock removal (Stock.stamp, src/stock.js:118-164) on gfx950.
//
// The heightfield and its side maps stay resident in HBM.  A batch of n stamps is applied
// with the reference's sequential semantics by giving every grid node one thread that
// visits, in stamp order, exactly the stamps whose bounding box covers it; nodes are
// independent, so the final H, err, wallD, lineCut and lineWall are bit-identical to n
// successive Stock.stamp calls, while each node's state is read from and written to HBM
// once per batch (the tile keeps it in registers).
//
//   stamp_prep   per stamp: bounding box (the reference's floor/ceil rule), radii and
//                rim height; per chunk of CHUNK stamps: union bounding box.
//   stamp_tiles  one 16x16 node tile per workgroup: walks the chunks whose box meets the
//                tile, compacts (in order, with wave ballots) the stamps of a chunk that
//                meet the tile and can still cut it into LDS, and lets each node thread
//                apply them in order.  A stamp can cut a node only where H > zmin (its
//                lowest reach), so stamps whose zmin is at or above the tile's highest node
//                are dropped, and each wave skips the stamps at or above its own highest
//                node: exact, since those stamps would change nothing.
//
// Removed volume per stamp (the value stamp() returns) is summed with one FP64 atomic per
// wavefront and stamp, so its last bits depend on the summation order.
#include <hip/hip_runtime.h>

#include "../../include/mcsim.h"

namespace mcs {

constexpr int TILE = 16;      // 16 x 16 nodes per workgroup
constexpr int CHUNK = 256;    // stamps compacted into LDS at a time
constexpr double EDGE_TIE = 2 * 0.0015;  // src/stock.js:6

struct StampDev {
  double cx, cy, tipZ, R, R2, tie, Rw2, zRim;
  double zmin;   // lowest z the stamp can cut to: both of its updates need H above it
  float errVal;
  int32_t line, type, i0, i1, j0, j1, index;
};

struct ChunkBox {
  int32_t i0, i1, j0, j1;
};

__global__ void stamp_prep(const mc_stamp* __restrict__ in, int n, double hx, double hy, int nx,
                           int ny, double zTop, double band, StampDev* __restrict__ out,
                           ChunkBox* __restrict__ box) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ int bi0[256], bi1[256], bj0[256], bj1[256];
  int i0 = 1 << 30, i1 = -1, j0 = 1 << 30, j1 = -1;
  if (s < n) {
    const mc_stamp st = in[s];
    StampDev d;
    d.cx = st.cx; d.cy = st.cy; d.tipZ = st.tipZ; d.R = st.R;
    d.type = st.type; d.line = st.line; d.errVal = (float)st.errVal; d.index = s;
    const double m = 2 * (hx > hy ? hx : hy);
    const double Rw = st.R + (m > band ? m : band);
    d.Rw2 = Rw * Rw;
    d.R2 = st.R * st.R;
    d.tie = d.R2 - EDGE_TIE * st.R;
    d.zRim = st.tipZ + (st.type == MC_TOOL_BALL ? st.R : 0);
    // a ball cuts to z = (tipZ + R) - sqrt(R^2 - r2) >= (tipZ + R) - R, a flat one to tipZ; the
    // wall update needs H > zRim >= zmin too, so a node with H <= zmin is left as it is
    d.zmin = !(st.R >= 0) ? -__builtin_inf() : st.type == MC_TOOL_BALL ? (st.tipZ + st.R) - st.R : st.tipZ;
    const double a = __builtin_floor((st.cx - Rw) / hx), b = __builtin_ceil((st.cx + Rw) / hx);
    const double c = __builtin_floor((st.cy - Rw) / hy), e = __builtin_ceil((st.cy + Rw) / hy);
    const bool live = st.tipZ < zTop && a <= nx - 1 && c <= ny - 1 && b >= 0 && e >= 0;
    d.i0 = live ? (int)(a > 0 ? a : 0) : 1;
    d.i1 = live ? (int)(b < nx - 1 ? b : nx - 1) : 0;
    d.j0 = live ? (int)(c > 0 ? c : 0) : 1;
    d.j1 = live ? (int)(e < ny - 1 ? e : ny - 1) : 0;
    out[s] = d;
    if (d.i0 <= d.i1 && d.j0 <= d.j1) { i0 = d.i0; i1 = d.i1; j0 = d.j0; j1 = d.j1; }
  }
  bi0[threadIdx.x] = i0; bi1[threadIdx.x] = i1; bj0[threadIdx.x] = j0; bj1[threadIdx.x] = j1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int t = threadIdx.x;
      bi0[t] = min(bi0[t], bi0[t + w]); bi1[t] = max(bi1[t], bi1[t + w]);
      bj0[t] = min(bj0[t], bj0[t + w]); bj1[t] = max(bj1[t], bj1[t + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) box[blockIdx.x] = ChunkBox{bi0[0], bi1[0], bj0[0], bj1[0]};
}

__global__ __launch_bounds__(TILE * TILE) void stamp_tiles(
    const StampDev* __restrict__ st, const ChunkBox* __restrict__ box, int n, int nx, int ny,
    double hx, double hy, double zB, float* __restrict__ H, float* __restrict__ err,
    float* __restrict__ wallD, int32_t* __restrict__ lineCut, int32_t* __restrict__ lineWall,
    double* __restrict__ vol) {
  __shared__ StampDev list[CHUNK];
  const int t = threadIdx.x;
  const int ti0 = blockIdx.x * TILE, tj0 = blockIdx.y * TILE;
  const int ti1 = min(ti0 + TILE, nx) - 1, tj1 = min(tj0 + TILE, ny) - 1;
  const int i = ti0 + (t % TILE), j = tj0 + (t / TILE);
  const bool own = i <= ti1 && j <= tj1;
  const int k = j * nx + i;
  float h = 0, e = 0, wd = 0;
  int lc = 0, lw = 0;
  if (own) { h = H[k]; e = err[k]; wd = wallD[k]; lc = lineCut[k]; lw = lineWall[k]; }
  const double xi = i * hx, yj = j * hy;
  const int nchunks = (n + CHUNK - 1) / CHUNK;
  const int lane = t & 63, wv = t >> 6;
  __shared__ int wcount[TILE * TILE / 64];
  __shared__ float wmaxs[TILE * TILE / 64];
  for (int c = 0; c < nchunks; c++) {
    const ChunkBox b = box[c];
    if (b.i1 < ti0 || b.i0 > ti1 || b.j1 < tj0 || b.j0 > tj1) continue;  // block-uniform
    // highest node of this wave and of the tile (H only falls, so a stale maximum is a
    // safe upper bound for the rest of the chunk)
    float wmax = own ? h : -__builtin_inff();
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
    if (lane == 0) wmaxs[wv] = wmax;
    // order-preserving compaction of the chunk's stamps that meet this tile and reach
    // below its highest node
    const int s = c * CHUNK + t;
    StampDev d;
    bool cand = false;
    if (s < n) {
      d = st[s];
      cand = !(d.i1 < ti0 || d.i0 > ti1 || d.j1 < tj0 || d.j0 > tj1) && d.i0 <= d.i1 && d.j0 <= d.j1;
    }
    __syncthreads();
    float tmax = wmaxs[0];
    for (int v = 1; v < TILE * TILE / 64; v++) tmax = fmaxf(tmax, wmaxs[v]);
    const bool hit = cand && (double)tmax > d.zmin;
    const unsigned long long m = __builtin_amdgcn_ballot_w64(hit);
    if (lane == 0) wcount[wv] = __builtin_popcountll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int v = 0; v < TILE * TILE / 64; v++) {
      before += v < wv ? wcount[v] : 0;
      total += wcount[v];
    }
    if (hit) list[before + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0))] = d;
    __syncthreads();
    for (int q = 0; q < total; q++) {
      const StampDev& p = list[q];
      if ((double)wmax <= p.zmin) continue;   // wave-uniform: no node of this wave can change
      double removed = 0;
      if (own && i >= p.i0 && i <= p.i1 && j >= p.j0 && j <= p.j1) {
        const double dx = xi - p.cx, dy = yj - p.cy;
        const double r2 = dx * dx + dy * dy;
        if (r2 >= p.tie) {
          if (r2 <= p.Rw2 && (double)h > p.zRim) {
            const double dd = __builtin_sqrt(r2) - p.R;
            if (dd < (double)wd) { wd = (float)dd; lw = p.line; }
          }
        } else {
          const double z = p.type == MC_TOOL_BALL ? p.tipZ + p.R - __builtin_sqrt(p.R2 - r2) : p.tipZ;
          if ((double)h > z) {
            removed = (double)h - z;
            h = (float)(z > zB ? z : zB);
            e = p.errVal;
            wd = __builtin_inff();
            lc = p.line;
          }
        }
      }
      if (vol && __any(removed != 0.0)) {
        double r = removed;
        for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
        if (lane == 0 && r != 0.0) atomicAdd(&vol[p.index], r);
      }
    }
    __syncthreads();
  }
  if (own) { H[k] = h; err[k] = e; wallD[k] = wd; lineCut[k] = lc; lineWall[k] = lw; }
}

}  // namespace mcs

namespace mcs {

// Stock.sample (src/stock.js:85-92): bilinear height, -inf outside the stock
__global__ void stock_sample(const float* __restrict__ H, int nx, int ny, double hx, double hy,
                             const double* __restrict__ xy, int n, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double fx = xy[2 * q] / hx, fy = xy[2 * q + 1] / hy;
  if (fx < 0 || fy < 0 || fx > nx - 1 || fy > ny - 1) { out[q] = -__builtin_inf(); return; }
  int i = (int)fx, j = (int)fy;
  if (i > nx - 2) i = nx - 2;
  if (j > ny - 2) j = ny - 2;
  const double tx = fx - i, ty = fy - j;
  const double a = H[j * nx + i], b = H[j * nx + i + 1], c = H[(j + 1) * nx + i],
               d = H[(j + 1) * nx + i + 1];
  out[q] = (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}

__global__ void fill_f32(float* p, size_t n, float v) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = v;
}

}  // namespace mcs

// ---- host-side launchers used by api.cpp
size_t mcs_stampdev_bytes() { return sizeof(mcs::StampDev); }
size_t mcs_chunkbox_bytes() { return sizeof(mcs::ChunkBox); }
int mcs_chunk() { return mcs::CHUNK; }

int mcs_launch_stamps(const mc_stamp* d_in, int n, void* d_stampdev, void* d_box, int nx, int ny,
                      double hx, double hy, double zTop, double zB, double band, float* H,
                      float* err, float* wallD, int32_t* lineCut, int32_t* lineWall, double* vol,
                      hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
  if (n <= 0) return 0;
  const int chunks = (n + mcs::CHUNK - 1) / mcs::CHUNK;
  hipLaunchKernelGGL(mcs::stamp_prep, dim3(chunks), dim3(mcs::CHUNK), 0, stream, d_in, n, hx, hy,
                     nx, ny, zTop, band, (mcs::StampDev*)d_stampdev, (mcs::ChunkBox*)d_box);
  const dim3 grid((nx + mcs::TILE - 1) / mcs::TILE, (ny + mcs::TILE - 1) / mcs::TILE);
  if (ev0 && hipEventRecord(ev0, stream) != hipSuccess) return -2;
  hipLaunchKernelGGL(mcs::stamp_tiles, grid, dim3(mcs::TILE * mcs::TILE), 0, stream,
                     (const mcs::StampDev*)d_stampdev, (const mcs::ChunkBox*)d_box, n, nx, ny, hx,
                     hy, zB, H, err, wallD, lineCut, lineWall, vol);
  if (ev1 && hipEventRecord(ev1, stream) != hipSuccess) return -2;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_sample(const float* H, int nx, int ny, double hx, double hy, const double* d_xy,
                      int n, double* d_out, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(mcs::stock_sample, dim3((n + 255) / 256), dim3(256), 0, stream, H, nx, ny,
                     hx, hy, d_xy, n, d_out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mcs_launch_fill(float* p, size_t n, float v, hipStream_t stream) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(mcs::fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
