// mlp.hip — the learned bracket surrogate: an MLP 24 -> 256 -> 256 -> 256 -> 1 (GELU, tanh
// form) that predicts log(alim) of a stability cell, trained and evaluated on the device.
//
// Every matrix product is one templated BF16 MFMA kernel (v_mfma_f32_16x16x32_bf16, FP32
// accumulate): 64x64 output tiles per 256-thread workgroup, 32-deep K steps staged in LDS,
// operands read K-contiguous or transposed through the LDS loader, optional split-K with an
// in-order reduction, and fused epilogues (bias + GELU storing the FP32 pre-activation and
// the BF16 activation; GELU backward producing the BF16 gradient of the pre-activation).
// Master weights, gradients and Adam moments are FP32; a training step (gather, forward,
// loss, backward, Adam, BF16 weight cast) is captured once into a hipGraph and replayed,
// with the step counter kept on the device.
//
// Quantisation points (the CPU oracle in oracle/mlp.py reproduces them): inputs, weights,
// activations H, the output gradient dY and the pre-activation gradients dZ are rounded to
// BF16 (round to nearest even) before they enter a product; everything else is FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mcsim.h"
#include "mc_host.h"

namespace mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int IN = MC_SURR_IN, INP = 32, HID = MC_SURR_HIDDEN;
constexpr int TM = 64, TN = 64, TK = 32, KP = TK + 8;  // LDS rows padded against bank conflicts
constexpr int NPARAM = IN * HID + HID + 2 * (HID * HID + HID) + HID + 1;
// offsets of the parameters in the public order: W1[24][256] b1 W2[256][256] b2 W3 b3 W4[256] b4
constexpr int OW1 = 0, OB1 = OW1 + IN * HID, OW2 = OB1 + HID, OB2 = OW2 + HID * HID,
              OW3 = OB2 + HID, OB3 = OW3 + HID * HID, OW4 = OB3 + HID, OB4 = OW4 + HID;
static_assert(OB4 + 1 == NPARAM, "parameter layout");

enum { EPI_F32 = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BWD = 3, EPI_SPLIT = 4 };

struct GemmArgs {
  const bf16* A; long lda;   // A_K: A(m,k) = A[m*lda + k]   else A(m,k) = A[k*lda + m]
  const bf16* B; long ldb;   // B_K: B(k,n) = B[n*ldb + k]   else B(k,n) = B[k*ldb + n]
  int M, N, K, kchunk;       // kchunk: K range of one split (multiple of TK)
  int avec, bvec;            // 16-byte vector loads allowed for A / B
  float* C; long ldc;        // FP32 result (EPI_F32, EPI_BIAS, EPI_BIAS_GELU: pre-activation)
  const float* bias;         // [N]
  bf16* H; long ldh;         // EPI_BIAS_GELU: activation; EPI_GELU_BWD: pre-activation gradient
  const float* Z; long ldz;  // EPI_GELU_BWD: pre-activation
};

__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f, u = c * (z + 0.044715f * z * z * z);
  return 0.5f * z * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f, u = c * (z + 0.044715f * z * z * z), t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

// stage a TM x TK (rows x k) operand tile into LDS rows of KP bf16, k contiguous
template <bool KCONTIG>
__device__ __forceinline__ void load_tile(const bf16* __restrict__ P, long ld, int rows, int K, int r0,
                                          int k0, bool vec, bf16 (*S)[KP]) {
  const int t = threadIdx.x;
  const bf16 zero = (bf16)0.0f;
  if (KCONTIG) {
    const int r = t >> 2, ks = (t & 3) * 8;
    const int row = r0 + r, k = k0 + ks;
    bf16x8 v;
    if (vec) {
      if (row < rows && k < K) v = *reinterpret_cast<const bf16x8*>(P + (long)row * ld + k);
      else for (int j = 0; j < 8; j++) v[j] = zero;
    } else {
      for (int j = 0; j < 8; j++) v[j] = (row < rows && k + j < K) ? P[(long)row * ld + k + j] : zero;
    }
    *reinterpret_cast<bf16x8*>(&S[r][ks]) = v;
  } else {
    const int kk = t >> 3, rs = (t & 7) * 8;
    const int k = k0 + kk, row = r0 + rs;
    bf16x8 v;
    if (vec) {
      if (k < K && row < rows) v = *reinterpret_cast<const bf16x8*>(P + (long)k * ld + row);
      else for (int j = 0; j < 8; j++) v[j] = zero;
    } else {
      for (int j = 0; j < 8; j++) v[j] = (k < K && row + j < rows) ? P[(long)k * ld + row + j] : zero;
    }
    for (int j = 0; j < 8; j++) S[rs + j][kk] = v[j];
  }
}

// C = A x B over one TM x TN tile; blockIdx.z selects the K split
template <bool AK, bool BK, int EPI>
__global__ __launch_bounds__(256) void gemm(GemmArgs g, float* __restrict__ ws) {
  __shared__ __attribute__((aligned(16))) bf16 As[TM][KP];
  __shared__ __attribute__((aligned(16))) bf16 Bs[TN][KP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int kb = blockIdx.z * g.kchunk, ke = min(g.K, kb + g.kchunk);
  f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += TK) {
    load_tile<AK>(g.A, g.lda, g.M, ke, m0, k0, g.avec, As);
    load_tile<BK>(g.B, g.ldb, g.N, ke, n0, k0, g.bvec, Bs);
    __syncthreads();
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + (lane & 15)][(lane >> 4) * 8]);
    for (int j = 0; j < 2; j++) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + j * 16 + (lane & 15)][(lane >> 4) * 8]);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // lane l holds rows 4*(l/16) + r, column l%16 of each 16 x 16 block
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r, n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const float v = acc[i][j][r];
        if (EPI == EPI_F32) {
          g.C[(long)m * g.ldc + n] = v;
        } else if (EPI == EPI_BIAS) {
          g.C[(long)m * g.ldc + n] = v + g.bias[n];
        } else if (EPI == EPI_BIAS_GELU) {
          const float z = v + g.bias[n];
          g.C[(long)m * g.ldc + n] = z;
          g.H[(long)m * g.ldh + n] = (bf16)gelu(z);
        } else if (EPI == EPI_GELU_BWD) {
          g.H[(long)m * g.ldh + n] = (bf16)(v * gelu_grad(g.Z[(long)m * g.ldz + n]));
        } else {
          ws[((long)blockIdx.z * g.M + m) * g.N + n] = v;
        }
      }
}

// C[m][n] = sum over splits s (in order) of ws[s][m][n]
__global__ void split_reduce(const float* __restrict__ ws, int splits, int M, int N, float* __restrict__ C, long ldc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  float s = 0;
  for (int z = 0; z < splits; z++) s += ws[(long)z * M * N + i];
  C[(i / N) * ldc + i % N] = s;
}

// column sums of a BF16 [rows][N] matrix: partial sums over row chunks, then in order
__global__ void colsum_part(const bf16* __restrict__ D, int rows, int N, int chunk, float* __restrict__ part) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (n >= N) return;
  float s = 0;
  const int r1 = min(rows, (c + 1) * chunk);
  for (int r = c * chunk; r < r1; r++) s += (float)D[(long)r * N + n];
  part[(long)c * N + n] = s;
}
__global__ void colsum_final(const float* __restrict__ part, int chunks, int N, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0;
  for (int c = 0; c < chunks; c++) s += part[(long)c * N + n];
  out[n] = s;
}

// minibatch rows: X (scaled FP32 [n][INP]) -> BF16 [B][INP], targets; the batch is
// perm[(step mod steps_per_epoch) * B + r] with the step counter read on the device
__global__ void gather(const float* __restrict__ X, const float* __restrict__ T, const int* __restrict__ perm,
                       const int* __restrict__ step, int spe, int B, bf16* __restrict__ Xb, float* __restrict__ Tb) {
  const int r = blockIdx.x, c = threadIdx.x;  // B blocks of INP threads
  const int base = (*step % spe) * B;
  const int src = perm ? perm[base + r] : r;
  Xb[(long)r * INP + c] = (bf16)X[(long)src * INP + c];
  if (c == 0) Tb[r] = T[src];
}

// loss = sum (y - t)^2 / (2B); dY = (y - t) / B in FP32 and BF16; db4 = sum dY.  One
// workgroup, fixed summation order.
__global__ __launch_bounds__(256) void loss_grad(const float* __restrict__ y, const float* __restrict__ t, int B,
                                                 float* __restrict__ dY, bf16* __restrict__ dYb,
                                                 float* __restrict__ db4, float* __restrict__ loss,
                                                 const int* __restrict__ step) {
  __shared__ float sl[256], sg[256];
  float l = 0, gsum = 0;
  for (int r = threadIdx.x; r < B; r += 256) {
    const float d = y[r] - t[r];
    const float gr = d / (float)B;
    dY[r] = gr;
    dYb[r] = (bf16)gr;
    l += d * d;
    gsum += gr;
  }
  sl[threadIdx.x] = l;
  sg[threadIdx.x] = gsum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) { sl[threadIdx.x] += sl[threadIdx.x + s]; sg[threadIdx.x] += sg[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *db4 = sg[0];
    loss[step ? *step : 0] = sl[0] / (2.0f * (float)B);
  }
}

// Adam with bias correction and a cosine learning-rate schedule over total steps
__global__ void adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M1,
                     float* __restrict__ M2, int n, const int* __restrict__ step, int total, float lr0,
                     float lr1, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = *step + 1;
  const float lr = lr1 + 0.5f * (lr0 - lr1) * (1.0f + cosf(3.14159265358979f * (float)min(t, total) / (float)total));
  const float g = G[i];
  const float m = b1 * M1[i] + (1.0f - b1) * g;
  const float v = b2 * M2[i] + (1.0f - b2) * g * g;
  M1[i] = m;
  M2[i] = v;
  const float mh = m / (1.0f - powf(b1, (float)t)), vh = v / (1.0f - powf(b2, (float)t));
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

// BF16 copies of the weights: W1 padded to INP rows, W2, W3, W4; thread 0 then advances
// the device step counter (this is the last kernel of a training step)
__global__ void cast_weights(const float* __restrict__ P, bf16* __restrict__ W1, bf16* __restrict__ W2,
                             bf16* __restrict__ W3, bf16* __restrict__ W4, int* __restrict__ step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < INP * HID) W1[i] = i < IN * HID ? (bf16)P[OW1 + i] : (bf16)0.0f;
  if (i < HID * HID) { W2[i] = (bf16)P[OW2 + i]; W3[i] = (bf16)P[OW3 + i]; }
  if (i < HID) W4[i] = (bf16)P[OW4 + i];
  if (step && i == 0) *step += 1;
}

__global__ void fill_i32(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace mlp

// ----------------------------------------------------------------------------- host side
using namespace mlp;

namespace {

struct Buf {
  void* p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return MC_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return mch::fail(MC_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    n = bytes;
    return MC_OK;
  }
  ~Buf() { if (p) (void)hipFree(p); }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

template <bool AK, bool BK, int EPI>
void launch_gemm(const GemmArgs& g0, int splits, float* ws, hipStream_t s) {
  GemmArgs g = g0;
  const int kt = (g.K + TK - 1) / TK;
  splits = std::max(1, std::min(splits, kt));
  g.kchunk = ((kt + splits - 1) / splits) * TK;
  splits = (g.K + g.kchunk - 1) / g.kchunk;
  dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, splits);
  if (EPI == EPI_SPLIT) {
    hipLaunchKernelGGL((gemm<AK, BK, EPI_SPLIT>), grid, dim3(256), 0, s, g, ws);
    const long mn = (long)g.M * g.N;
    hipLaunchKernelGGL(split_reduce, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, ws, splits, g.M, g.N, g.C, g.ldc);
  } else {
    hipLaunchKernelGGL((gemm<AK, BK, EPI>), grid, dim3(256), 0, s, g, ws);
  }
}

bool vec_ok(const void* p, long ld, int contig_extent) {
  return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && ld % 8 == 0 && contig_extent % 8 == 0;
}

}  // namespace

struct mc_mlp {
  float xm[IN], xs[IN], ym = 0, ys = 1;     // input / target scaling
  Buf P, G, M1, M2;                         // FP32 parameters, gradients, Adam moments
  Buf W1, W2, W3, W4;                       // BF16 weight copies
  int cap = 0;                              // rows the activation buffers hold
  Buf Xb, Tb, Z1, Z2, Z3, H1, H2, H3, Y, dY, dYb, dZa, dZb, ws, part, step, loss;
  int ws_floats = 0;

  int init_params(uint64_t seed) {
    std::vector<float> h(NPARAM, 0.0f);
    // uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)), from a splitmix64 stream
    uint64_t st = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
      const double a = std::sqrt(6.0 / (fan_in + fan_out));
      for (int i = 0; i < count; i++) h[off + i] = (float)((2 * next() - 1) * a);
    };
    fill(OW1, IN, HID, IN * HID);
    fill(OW2, HID, HID, HID * HID);
    fill(OW3, HID, HID, HID * HID);
    fill(OW4, HID, 1, HID);
    for (int i = 0; i < IN; i++) { xm[i] = 0; xs[i] = 1; }
    int st2;
    if ((st2 = P.ensure(NPARAM * 4)) || (st2 = G.ensure(NPARAM * 4)) || (st2 = M1.ensure(NPARAM * 4)) ||
        (st2 = M2.ensure(NPARAM * 4)) || (st2 = W1.ensure(INP * HID * 2)) || (st2 = W2.ensure(HID * HID * 2)) ||
        (st2 = W3.ensure(HID * HID * 2)) || (st2 = W4.ensure(HID * 2)) || (st2 = step.ensure(4)))
      return st2;
    hipStream_t s = mch::stream();
    MCH_HIPCHK(hipMemcpyAsync(P.p, h.data(), NPARAM * 4, hipMemcpyHostToDevice, s));
    return cast(s);
  }
  int cast(hipStream_t s) {
    hipLaunchKernelGGL(cast_weights, dim3(HID * HID / 256), dim3(256), 0, s, P.as<float>(), W1.as<bf16>(),
                       W2.as<bf16>(), W3.as<bf16>(), W4.as<bf16>(), (int*)nullptr);
    MCH_HIPCHK(hipGetLastError());
    return MC_OK;
  }
  int reserve(int rows) {
    if (rows <= cap) return MC_OK;
    int st;
    const size_t r = (size_t)rows;
    if ((st = Xb.ensure(r * INP * 2)) || (st = Tb.ensure(r * 4)) || (st = Z1.ensure(r * HID * 4)) ||
        (st = Z2.ensure(r * HID * 4)) || (st = Z3.ensure(r * HID * 4)) || (st = H1.ensure(r * HID * 2)) ||
        (st = H2.ensure(r * HID * 2)) || (st = H3.ensure(r * HID * 2)) || (st = Y.ensure(r * 4)) ||
        (st = dY.ensure(r * 4)) || (st = dYb.ensure(r * 2)) || (st = dZa.ensure(r * HID * 2)) ||
        (st = dZb.ensure(r * HID * 2)) ||
        (st = part.ensure(64 * HID * 4)))
      return st;
    // split-K workspace: up to 64 splits of a 256 x 256 product
    ws_floats = 64 * HID * HID;
    if ((st = ws.ensure((size_t)ws_floats * 4))) return st;
    cap = rows;
    return MC_OK;
  }

  // forward over B rows already in Xb: Z/H per layer, Y
  void forward(int B, hipStream_t s) {
    GemmArgs g{};
    g.M = B; g.N = HID;
    g.A = Xb.as<bf16>(); g.lda = INP; g.K = INP; g.avec = 1;
    g.B = W1.as<bf16>(); g.ldb = HID; g.bvec = 1;
    g.C = Z1.as<float>(); g.ldc = HID; g.bias = P.as<float>() + OB1; g.H = H1.as<bf16>(); g.ldh = HID;
    launch_gemm<true, false, EPI_BIAS_GELU>(g, 1, nullptr, s);
    g.A = H1.as<bf16>(); g.lda = HID; g.K = HID;
    g.B = W2.as<bf16>(); g.C = Z2.as<float>(); g.bias = P.as<float>() + OB2; g.H = H2.as<bf16>();
    launch_gemm<true, false, EPI_BIAS_GELU>(g, 1, nullptr, s);
    g.A = H2.as<bf16>(); g.B = W3.as<bf16>(); g.C = Z3.as<float>(); g.bias = P.as<float>() + OB3; g.H = H3.as<bf16>();
    launch_gemm<true, false, EPI_BIAS_GELU>(g, 1, nullptr, s);
    GemmArgs o{};
    o.M = B; o.N = 1; o.K = HID;
    o.A = H3.as<bf16>(); o.lda = HID; o.avec = 1;
    o.B = W4.as<bf16>(); o.ldb = 1; o.bvec = 0;   // B(k, 0) = W4[k]
    o.C = Y.as<float>(); o.ldc = 1; o.bias = P.as<float>() + OB4;
    launch_gemm<true, false, EPI_BIAS>(o, 1, nullptr, s);
  }

  // backward over B rows (targets in Tb): gradients into G in the public order; the
  // pre-activation gradients alternate between dZa (layer 3, layer 1) and dZb (layer 2)
  void backward(int B, hipStream_t s, const int* dstep) {
    float* Gp = G.as<float>();
    hipLaunchKernelGGL(loss_grad, dim3(1), dim3(256), 0, s, Y.as<float>(), Tb.as<float>(), B, dY.as<float>(),
                       dYb.as<bf16>(), Gp + OB4, loss.as<float>(), dstep);
    const int splits = std::max(1, std::min(64, B / 256));
    // dW4[k] = sum_b H3[b][k] dY[b]
    GemmArgs g{};
    g.M = HID; g.N = 1; g.K = B;
    g.A = H3.as<bf16>(); g.lda = HID; g.avec = vec_ok(g.A, g.lda, g.M);   // A(m=k, k=b) = H3[b][m]
    g.B = dYb.as<bf16>(); g.ldb = 1; g.bvec = 0;                          // B(k=b, 0) = dY[b]
    g.C = Gp + OW4; g.ldc = 1;
    launch_gemm<false, false, EPI_SPLIT>(g, splits, ws.as<float>(), s);
    // dZ3 = (dY x W4^T) * gelu'(Z3)
    bf16* dz = dZa.as<bf16>();
    GemmArgs d{};
    d.M = B; d.N = HID; d.K = 1;
    d.A = dYb.as<bf16>(); d.lda = 1; d.avec = 0;           // A(b, 0) = dY[b]
    d.B = W4.as<bf16>(); d.ldb = 1; d.bvec = 0;            // B(0, n) = W4[n]
    d.H = dz; d.ldh = HID; d.Z = Z3.as<float>(); d.ldz = HID;
    launch_gemm<true, true, EPI_GELU_BWD>(d, 1, nullptr, s);
    const bf16* Hprev[3] = {Xb.as<bf16>(), H1.as<bf16>(), H2.as<bf16>()};
    const float* Zprev[3] = {nullptr, Z1.as<float>(), Z2.as<float>()};
    const bf16* Wl[3] = {W1.as<bf16>(), W2.as<bf16>(), W3.as<bf16>()};
    const int OW[3] = {OW1, OW2, OW3}, OB[3] = {OB1, OB2, OB3};
    for (int l = 2; l >= 0; l--) {
      // db = column sums of dZ
      const int chunks = std::max(1, std::min(64, B / 64)), chunk = (B + chunks - 1) / chunks;
      hipLaunchKernelGGL(colsum_part, dim3(HID / 64, chunks), dim3(64), 0, s, dz, B, HID, chunk, part.as<float>());
      hipLaunchKernelGGL(colsum_final, dim3(HID / 64), dim3(64), 0, s, part.as<float>(), chunks, HID, Gp + OB[l]);
      // dW = Hprev^T x dZ (layer 1 has 24 real input rows)
      GemmArgs w{};
      w.M = l == 0 ? IN : HID; w.N = HID; w.K = B;
      w.A = Hprev[l]; w.lda = l == 0 ? INP : HID; w.avec = vec_ok(w.A, w.lda, w.M);  // A(m=i, k=b) = Hprev[b][i]
      w.B = dz; w.ldb = HID; w.bvec = vec_ok(w.B, w.ldb, w.N);                        // B(k=b, n) = dZ[b][n]
      w.C = Gp + OW[l]; w.ldc = HID;
      launch_gemm<false, false, EPI_SPLIT>(w, splits, ws.as<float>(), s);
      if (l == 0) break;
      // dZprev = (dZ x W^T) * gelu'(Zprev)
      bf16* nz = dz == dZa.as<bf16>() ? dZb.as<bf16>() : dZa.as<bf16>();
      GemmArgs d2{};
      d2.M = B; d2.N = HID; d2.K = HID;
      d2.A = dz; d2.lda = HID; d2.avec = 1;                  // A(b, k=o) = dZ[b][o]
      d2.B = Wl[l]; d2.ldb = HID; d2.bvec = 1;               // B(k=o, n=i) = W[i][o]
      d2.H = nz; d2.ldh = HID; d2.Z = Zprev[l]; d2.ldz = HID;
      launch_gemm<true, true, EPI_GELU_BWD>(d2, 1, nullptr, s);
      dz = nz;
    }
  }
};

namespace {

// scaled, padded FP32 rows [n][INP] and scaled targets
void scale_rows(const mc_mlp* m, const float* X, const float* t, int n, std::vector<float>& xs, std::vector<float>& ts) {
  xs.assign((size_t)n * INP, 0.0f);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < IN; c++) xs[(size_t)r * INP + c] = (X[(size_t)r * IN + c] - m->xm[c]) / m->xs[c];
  if (t) {
    ts.resize(n);
    for (int r = 0; r < n; r++) ts[r] = (t[r] - m->ym) / m->ys;
  }
}

int check_rows(const float* X, int n) {
  for (size_t i = 0; i < (size_t)n * IN; i++)
    if (!std::isfinite(X[i])) return mch::fail(MC_ERR_ARG, "feature %zu of row %zu is not finite", i % IN, i / IN);
  return MC_OK;
}

}  // namespace

extern "C" {

int mc_mlp_nparams(void) { return NPARAM; }

int mc_mlp_create(uint64_t seed, mc_mlp** out) {
  int st = mch::init();
  if (st) return st;
  if (!out) return mch::fail(MC_ERR_ARG, "null output handle");
  std::lock_guard<std::mutex> lk(mch::lock());
  mc_mlp* m = new mc_mlp();
  if ((st = m->init_params(seed)) || (st = (hipStreamSynchronize(mch::stream()) == hipSuccess ? 0 : MC_ERR_HIP))) {
    delete m;
    return st == MC_ERR_HIP ? mch::fail(MC_ERR_HIP, "parameter upload failed") : st;
  }
  *out = m;
  return MC_OK;
}

int mc_mlp_destroy(mc_mlp* m) {
  if (!m) return MC_OK;
  std::lock_guard<std::mutex> lk(mch::lock());
  (void)hipStreamSynchronize(mch::stream());
  delete m;
  return MC_OK;
}

int mc_mlp_set_scaling(mc_mlp* m, const float* x_mean, const float* x_std, float y_mean, float y_std) {
  if (!m || !x_mean || !x_std) return mch::fail(MC_ERR_ARG, "null argument");
  for (int i = 0; i < IN; i++)
    if (!(x_std[i] > 0) || !std::isfinite(x_mean[i])) return mch::fail(MC_ERR_ARG, "feature %d: need std > 0", i);
  if (!(y_std > 0) || !std::isfinite(y_mean)) return mch::fail(MC_ERR_ARG, "need target std > 0");
  for (int i = 0; i < IN; i++) { m->xm[i] = x_mean[i]; m->xs[i] = x_std[i]; }
  m->ym = y_mean;
  m->ys = y_std;
  return MC_OK;
}

int mc_mlp_get_scaling(const mc_mlp* m, float* x_mean, float* x_std, float* y_mean, float* y_std) {
  if (!m || !x_mean || !x_std || !y_mean || !y_std) return mch::fail(MC_ERR_ARG, "null argument");
  for (int i = 0; i < IN; i++) { x_mean[i] = m->xm[i]; x_std[i] = m->xs[i]; }
  *y_mean = m->ym;
  *y_std = m->ys;
  return MC_OK;
}

int mc_mlp_params(mc_mlp* m, float* params, int write) {
  if (!m || !params) return mch::fail(MC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(mch::lock());
  hipStream_t s = mch::stream();
  if (write) {
    for (int i = 0; i < NPARAM; i++)
      if (!std::isfinite(params[i])) return mch::fail(MC_ERR_ARG, "parameter %d is not finite", i);
    MCH_HIPCHK(hipMemcpyAsync(m->P.p, params, NPARAM * 4, hipMemcpyHostToDevice, s));
    int st = m->cast(s);
    if (st) return st;
  } else {
    MCH_HIPCHK(hipMemcpyAsync(params, m->P.p, NPARAM * 4, hipMemcpyDeviceToHost, s));
  }
  MCH_HIPCHK(hipStreamSynchronize(s));
  return MC_OK;
}

int mc_mlp_grads(mc_mlp* m, const float* X, const float* t, int n, float* grads, float* loss) {
  if (!m || !X || !t || n <= 0 || !grads) return mch::fail(MC_ERR_ARG, "bad arguments");
  int st = check_rows(X, n);
  if (st) return st;
  std::lock_guard<std::mutex> lk(mch::lock());
  hipStream_t s = mch::stream();
  if ((st = m->reserve(n)) || (st = m->loss.ensure(4))) return st;
  std::vector<float> xs, ts;
  scale_rows(m, X, t, n, xs, ts);
  Buf dx, dt;
  if ((st = dx.ensure(xs.size() * 4)) || (st = dt.ensure(ts.size() * 4))) return st;
  MCH_HIPCHK(hipMemcpyAsync(dx.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, s));
  MCH_HIPCHK(hipMemcpyAsync(dt.p, ts.data(), ts.size() * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, m->step.as<int>(), 1, 0);
  hipLaunchKernelGGL(gather, dim3(n), dim3(INP), 0, s, dx.as<float>(), dt.as<float>(), (const int*)nullptr,
                     m->step.as<int>(), 1, n, m->Xb.as<bf16>(), m->Tb.as<float>());
  m->forward(n, s);
  m->backward(n, s, nullptr);
  MCH_HIPCHK(hipGetLastError());
  MCH_HIPCHK(hipMemcpyAsync(grads, m->G.p, NPARAM * 4, hipMemcpyDeviceToHost, s));
  float l = 0;
  MCH_HIPCHK(hipMemcpyAsync(&l, m->loss.p, 4, hipMemcpyDeviceToHost, s));
  MCH_HIPCHK(hipStreamSynchronize(s));
  if (loss) *loss = l;
  return MC_OK;
}

int mc_mlp_train(mc_mlp* m, const float* X, const float* t, int n, const mc_train_opts* o, float* epoch_loss) {
  if (!m || !X || !t || !o) return mch::fail(MC_ERR_ARG, "null argument");
  if (o->batch < 64 || o->batch % 64 || o->batch > n || o->epochs < 1)
    return mch::fail(MC_ERR_ARG, "need 64 <= batch <= n, batch a multiple of 64, epochs >= 1");
  if (!(o->lr > 0) || !(o->lr_final >= 0) || !(o->beta1 >= 0 && o->beta1 < 1) || !(o->beta2 >= 0 && o->beta2 < 1) || !(o->eps > 0))
    return mch::fail(MC_ERR_ARG, "bad Adam settings");
  int st = check_rows(X, n);
  if (st) return st;
  for (int r = 0; r < n; r++)
    if (!std::isfinite(t[r])) return mch::fail(MC_ERR_ARG, "target %d is not finite", r);
  if (o->fit_scaling) {
    // per-feature mean / std of the training rows (constant features keep std 1)
    double ym = 0, yv = 0;
    for (int c = 0; c < IN; c++) {
      double mu = 0, var = 0;
      for (int r = 0; r < n; r++) mu += X[(size_t)r * IN + c];
      mu /= n;
      for (int r = 0; r < n; r++) { const double d = X[(size_t)r * IN + c] - mu; var += d * d; }
      var /= n;
      m->xm[c] = (float)mu;
      m->xs[c] = var > 1e-12 ? (float)std::sqrt(var) : 1.0f;
    }
    for (int r = 0; r < n; r++) ym += t[r];
    ym /= n;
    for (int r = 0; r < n; r++) yv += (t[r] - ym) * (t[r] - ym);
    yv /= n;
    m->ym = (float)ym;
    m->ys = yv > 1e-12 ? (float)std::sqrt(yv) : 1.0f;
  }
  std::lock_guard<std::mutex> lk(mch::lock());
  hipStream_t s = mch::stream();
  const int B = o->batch, spe = n / B, total = spe * o->epochs;
  if ((st = m->reserve(B)) || (st = m->loss.ensure((size_t)total * 4))) return st;
  std::vector<float> xs, ts;
  scale_rows(m, X, t, n, xs, ts);
  Buf dx, dt, dp;
  if ((st = dx.ensure(xs.size() * 4)) || (st = dt.ensure(ts.size() * 4)) || (st = dp.ensure((size_t)n * 4))) return st;
  MCH_HIPCHK(hipMemcpyAsync(dx.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, s));
  MCH_HIPCHK(hipMemcpyAsync(dt.p, ts.data(), ts.size() * 4, hipMemcpyHostToDevice, s));
  // fresh optimiser state
  std::vector<float> zeros(NPARAM, 0.0f);
  MCH_HIPCHK(hipMemcpyAsync(m->M1.p, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice, s));
  MCH_HIPCHK(hipMemcpyAsync(m->M2.p, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, m->step.as<int>(), 1, 0);
  MCH_HIPCHK(hipStreamSynchronize(s));

  // one training step (the step counter lives on the device, so every launch of the sequence
  // is the next step): captured once into a graph and replayed, or, with MC_MLP_EAGER set in
  // the environment, launched kernel by kernel (the same kernels and results; for kernel
  // tracers, which then see ordinary dispatches)
  auto step_kernels = [&]() {
    hipLaunchKernelGGL(gather, dim3(B), dim3(INP), 0, s, dx.as<float>(), dt.as<float>(), dp.as<int>(),
                       m->step.as<int>(), spe, B, m->Xb.as<bf16>(), m->Tb.as<float>());
    m->forward(B, s);
    m->backward(B, s, m->step.as<int>());
    hipLaunchKernelGGL(adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, m->P.as<float>(), m->G.as<float>(),
                       m->M1.as<float>(), m->M2.as<float>(), NPARAM, m->step.as<int>(), total, o->lr, o->lr_final,
                       o->beta1, o->beta2, o->eps);
    hipLaunchKernelGGL(cast_weights, dim3(HID * HID / 256), dim3(256), 0, s, m->P.as<float>(), m->W1.as<bf16>(),
                       m->W2.as<bf16>(), m->W3.as<bf16>(), m->W4.as<bf16>(), m->step.as<int>());
  };
  const bool eager = getenv("MC_MLP_EAGER") != nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (!eager) {
    MCH_HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    step_kernels();
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) return mch::fail(MC_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ce != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return mch::fail(MC_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ce));
    }
  }
  // epochs: a seeded Fisher-Yates permutation per epoch, then spe replays
  uint64_t rs = o->seed * 0x9E3779B97F4A7C15ull + 1;
  auto rnd = [&]() { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; };
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  hipEvent_t e0, e1;
  MCH_HIPCHK(hipEventCreate(&e0));
  MCH_HIPCHK(hipEventCreate(&e1));
  MCH_HIPCHK(hipEventRecord(e0, s));
  int err = MC_OK;
  for (int ep = 0; ep < o->epochs && !err; ep++) {
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
    if (hipMemcpyAsync(dp.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess) err = MC_ERR_HIP;
    // the copy must land before the replays read it; the host array is reused next epoch
    if (!err && hipStreamSynchronize(s) != hipSuccess) err = MC_ERR_HIP;
    for (int k = 0; k < spe && !err; k++) {
      if (eager) {
        step_kernels();
        if (hipGetLastError() != hipSuccess) err = MC_ERR_HIP;
      } else if (hipGraphLaunch(exec, s) != hipSuccess) {
        err = MC_ERR_HIP;
      }
    }
  }
  (void)hipEventRecord(e1, s);
  std::vector<float> hl(total);
  if (!err && hipMemcpyAsync(hl.data(), m->loss.p, (size_t)total * 4, hipMemcpyDeviceToHost, s) != hipSuccess) err = MC_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) err = MC_ERR_HIP;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  mch::set_last_ms(ms);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (err) return mch::fail(MC_ERR_HIP, "training step failed: %s", hipGetErrorString(hipGetLastError()));
  if (epoch_loss)
    for (int ep = 0; ep < o->epochs; ep++) {
      double a = 0;
      for (int k = 0; k < spe; k++) a += hl[(size_t)ep * spe + k];
      epoch_loss[ep] = (float)(a / spe);
    }
  return MC_OK;
}

int mc_mlp_predict(mc_mlp* m, const float* X, int n, float* out) {
  if (!m || n < 0 || (n > 0 && (!X || !out))) return mch::fail(MC_ERR_ARG, "bad arguments");
  if (n == 0) return MC_OK;
  int st = check_rows(X, n);
  if (st) return st;
  std::lock_guard<std::mutex> lk(mch::lock());
  hipStream_t s = mch::stream();
  const int CH = 16384;
  const int rows = std::min(n, CH);
  if ((st = m->reserve(rows))) return st;
  std::vector<float> xs, ts;
  scale_rows(m, X, nullptr, n, xs, ts);
  Buf dx;
  if ((st = dx.ensure(xs.size() * 4))) return st;
  MCH_HIPCHK(hipMemcpyAsync(dx.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, m->step.as<int>(), 1, 0);
  std::vector<float> y(n);
  for (int r0 = 0; r0 < n; r0 += CH) {
    const int b = std::min(CH, n - r0);
    // rows r0.. of the scaled input (no permutation; targets unused)
    hipLaunchKernelGGL(gather, dim3(b), dim3(INP), 0, s, dx.as<float>() + (size_t)r0 * INP, dx.as<float>(),
                       (const int*)nullptr, m->step.as<int>(), 1, b, m->Xb.as<bf16>(), m->Tb.as<float>());
    m->forward(b, s);
    MCH_HIPCHK(hipMemcpyAsync(y.data() + r0, m->Y.p, (size_t)b * 4, hipMemcpyDeviceToHost, s));
  }
  MCH_HIPCHK(hipGetLastError());
  MCH_HIPCHK(hipStreamSynchronize(s));
  for (int r = 0; r < n; r++) out[r] = y[r] * m->ys + m->ym;
  return MC_OK;
}

// The 24 surrogate inputs of a cell (raw; mc_mlp_train fits their scaling).  Stability lobes
// of a mode repeat with its phase eps = 2 pi fn / f_tooth (f_tooth = rpm * flutes / 60), so
// each mode enters through its stiffness, damping ratio and that phase:
//   0-15  x modes 1-2, y modes 1-2: log10 k, zeta, sin eps, cos eps (zeros for an absent mode)
//   16    helix angle [deg]   17 flutes   18 diameter [mm]   19 immersion: arc / pi
//   20    log10 rpm   21 revolutions per trial
//   22    ln of the analytic (zero-order) limit at this speed, mc_analytic_limits, clipped
//         to [1e-3, 1e3] mm: the lobe structure of the linear model, which the network
//         corrects for process damping, the size effect and the time-domain verdict
//   23    log10 fn of the most flexible mode (lowest k over both axes)
// The cutting coefficients enter only through the analytic limit: the training
// distribution (the lobe grid's) holds them fixed.
int mc_surrogate_features(const mc_cell* cells, int n, float* out) {
  if (n < 0 || (n > 0 && (!cells || !out))) return mch::fail(MC_ERR_ARG, "bad arguments");
  std::vector<double> zoa(n);
  int st = mc_analytic_limits(cells, n, zoa.data());
  if (st) return st;
  for (int i = 0; i < n; i++) {
    const mc_cell& c = cells[i];
    float* f = out + (size_t)i * IN;
    for (int j = 0; j < IN; j++) f[j] = 0.0f;
    const double ftooth = (c.S > 100 ? c.S : 100.0) * c.tool.N / 60.0;
    for (int a = 0; a < 2; a++) {
      const mc_mode* md = a == 0 ? c.mx : c.my;
      const int nm = a == 0 ? c.nmx : c.nmy;
      for (int j = 0; j < 2 && j < nm; j++) {
        if (!(md[j].fn > 0) || !(md[j].k > 0)) return mch::fail(MC_ERR_ARG, "cell %d: mode needs fn > 0, k > 0", i);
        const double eps = 2 * 3.141592653589793 * (md[j].fn / ftooth);
        float* m = f + a * 8 + j * 4;
        m[0] = (float)std::log10(md[j].k);
        m[1] = (float)md[j].zeta;
        m[2] = (float)std::sin(eps);
        m[3] = (float)std::cos(eps);
      }
    }
    double kmin = INFINITY, fsoft = 1;
    for (int j = 0; j < c.nmx && j < MC_MAX_MODES; j++) if (c.mx[j].k < kmin) { kmin = c.mx[j].k; fsoft = c.mx[j].fn; }
    for (int j = 0; j < c.nmy && j < MC_MAX_MODES; j++) if (c.my[j].k < kmin) { kmin = c.my[j].k; fsoft = c.my[j].fn; }
    const double za = zoa[i] < 1e-3 ? 1e-3 : zoa[i] > 1e3 || !(zoa[i] == zoa[i]) ? 1e3 : zoa[i];
    f[22] = (float)std::log(za);
    f[23] = (float)std::log10(fsoft);
    f[16] = (float)c.tool.helixDeg;
    f[17] = (float)c.tool.N;
    f[18] = (float)(2 * c.tool.R);
    f[19] = (float)((c.phiEx - c.phiSt) / 3.141592653589793);
    f[20] = (float)std::log10(c.S > 100 ? c.S : 100.0);
    f[21] = (float)c.revs;
    for (int j = 0; j < IN; j++)
      if (!std::isfinite(f[j])) return mch::fail(MC_ERR_ARG, "cell %d: feature %d is not finite", i, j);
  }
  return MC_OK;
}

}  // extern "C"
