// Learned preconditioner: a 3-layer message-passing network over the bond
// graph, BF16 MFMA with FP32 accumulation, hand-written forward, backward and
// Adam kernels. No counterpart in the reference.
//
// Network (per node i, mean aggregation over the bond list N(i)):
//   x_i   = [r_i / s_r, (D^-1 r)_i / s_z, kinematic flag, deg_i / 128]  (8 features)
//   h1_i  = tanh([x_i,  mean_j x_j ] W1 + b1)   width 32
//   h2_i  = tanh([h1_i, mean_j h1_j] W2 + b2)   width 32
//   o_i   =      [h2_i, mean_j h2_j] W3 + b3    3 outputs
//   z     = D^-1 r + s_z o   (zero on kinematic DOFs)
// D^-1 is the 3x3 block-Jacobi inverse; s_r, s_z are the RMS of r and D^-1 r,
// so z(a r) = a z(r). Flexible CG absorbs the nonlinearity of z(r).
//
// Each layer is one kernel per 64-node tile (4 waves): the [self, mean]
// input rows are built in LDS as BF16, the weights sit in LDS, and each wave
// computes a 16-row slab with v_mfma_f32_16x16x16 BF16 (FP32 accumulate).
// Backward per layer: gz = g * (1 - h^2); per-tile partials of dW = X^T gz and
// db by MFMA, reduced deterministically afterwards; gX = gz W^T by MFMA; the
// aggregation's transpose is a gather (symmetric bond lists):
//   g_in_j = gX_self_j + sum_{i in N(j)} gX_agg_i / deg_i.
// Parameter layout (FP32 master copy, per layer): Ws[din][dout], Wn[din][dout],
// b[dout], layers [8->32, 32->32, 32->3].

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "pdg_ctx.hpp"
#include "pdg_device.hpp"

namespace pdg {

namespace gnn {

constexpr int NL = 3;
constexpr int DIN[NL] = {8, 32, 32};     // logical
constexpr int DOUT[NL] = {32, 32, 3};
constexpr int DINP[NL] = {16, 32, 32};   // padded to the MFMA K granularity
constexpr int DOUTP[NL] = {32, 32, 16};
constexpr int TILE = 64;

inline int layer_params(int l) { return 2 * DIN[l] * DOUT[l] + DOUT[l]; }
inline int total_params() { return layer_params(0) + layer_params(1) + layer_params(2); }
inline int param_off(int l) {
  int o = 0;
  for (int k = 0; k < l; k++) o += layer_params(k);
  return o;
}

}  // namespace gnn

struct GnnState {
  int n = 0;
  int ntiles = 0;
  bool trained = false;
  float* params = nullptr;   // master FP32
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  unsigned short* wt[gnn::NL] = {};  // forward weights, [DOUTP][2 DINP] bf16
  unsigned short* wn[gnn::NL] = {};  // backward weights, [2 DINP][DOUTP] bf16
  float* bias[gnn::NL] = {};         // [DOUTP]
  unsigned short* h[gnn::NL + 1] = {};  // activations: h[0] = features (N x 16), h[1], h[2] (N x 32)
  float* out = nullptr;       // N x 3
  float* gbuf[2] = {};        // N x 32 fp32 gradients
  float* gagg = nullptr;      // N x 32
  float* part = nullptr;      // per-tile partial dW/db
  float* scales = nullptr;    // [s_r, s_z, loss]
  double* z0 = nullptr;       // D^-1 r (3N)
  double* dz = nullptr;       // z - e (3N), training
  double* kdz = nullptr;      // K (z - e) (3N), training
  double* red = nullptr;      // reduction partials
  int nred = 0;
  int step = 0;
};

namespace {

using namespace dev;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// C(16x16) += A(16 x K) * B(K x 16); A row-major (lda), Bt = B^T row-major
// (ldb), both bf16 in LDS. Lane l holds C[4*(l/16) + i][l%16], i = 0..3.
template <int K>
__device__ __forceinline__ f32x4 mfma_block(const unsigned short* A, int lda,
                                            const unsigned short* Bt, int ldb, f32x4 acc) {
  const int l = threadIdx.x & 63;
  const int r = l & 15, kg = (l >> 4) * 4;
#pragma unroll
  for (int ks = 0; ks < K; ks += 16) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(A + r * lda + ks + kg);
    const bf16x4 b = *reinterpret_cast<const bf16x4*>(Bt + r * ldb + ks + kg);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// Builds the [self, mean over neighbours] bf16 input rows of a 64-node tile.
template <int D>
__device__ void build_inputs(int n, const int* __restrict__ row, const int* __restrict__ col,
                             const unsigned short* __restrict__ hin, int node0,
                             unsigned short* X, int ldx) {
  // self part
  for (int q = threadIdx.x; q < gnn::TILE * D; q += 256) {
    const int r = q / D, c = q - r * D;
    const int node = node0 + r;
    X[r * ldx + c] = node < n ? hin[(size_t)node * D + c] : (unsigned short)0;
  }
  // mean over neighbours: 4 threads per node, D/4 features each, read as one
  // 8- or 16-byte vector per neighbour; four neighbours are loaded per round
  // (independent partial sums, combined in a fixed order) so their loads
  // overlap
  constexpr int CH = D / 4;
  constexpr int NW = CH / 2;  // 32-bit words (two bf16 each) per chunk
  const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int node = node0 + r;
  float acc[4][CH];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int c = 0; c < CH; c++) acc[u][c] = 0.0f;
  int deg = 0;
  auto add_row = [&](float (&a)[CH], int j) {
    const unsigned* w = reinterpret_cast<const unsigned*>(hin + (size_t)j * D + q * CH);
    unsigned v[NW];
#pragma unroll
    for (int k = 0; k < NW; k++) v[k] = w[k];
#pragma unroll
    for (int k = 0; k < NW; k++) {
      a[2 * k] += __uint_as_float(v[k] << 16);
      a[2 * k + 1] += __uint_as_float(v[k] & 0xffff0000u);
    }
  };
  if (node < n) {
    const int rs = row[node], re = row[node + 1];
    deg = re - rs;
    int b = rs;
    for (; b + 4 <= re; b += 4) {
      const int j0 = col[b], j1 = col[b + 1], j2 = col[b + 2], j3 = col[b + 3];
      add_row(acc[0], j0);
      add_row(acc[1], j1);
      add_row(acc[2], j2);
      add_row(acc[3], j3);
    }
    for (; b < re; b++) add_row(acc[0], col[b]);
  }
#pragma unroll
  for (int c = 0; c < CH; c++) acc[0][c] = (acc[0][c] + acc[1][c]) + (acc[2][c] + acc[3][c]);
  const float inv = deg > 0 ? 1.0f / (float)deg : 0.0f;
#pragma unroll
  for (int c = 0; c < CH; c++) X[r * ldx + D + q * CH + c] = f2bf(acc[0][c] * inv);
}

// Forward of one layer over 64-node tiles.
template <int DI, int DO, bool ACT>
__global__ __launch_bounds__(256) void k_fwd(int n, const int* __restrict__ row,
                                             const int* __restrict__ col,
                                             const unsigned short* __restrict__ hin,
                                             const unsigned short* __restrict__ wt,
                                             const float* __restrict__ bias,
                                             unsigned short* __restrict__ hout,
                                             float* __restrict__ out3) {
  constexpr int K = 2 * DI;
  constexpr int LD = K + 4;
  __shared__ __align__(16) unsigned short X[gnn::TILE * LD];
  __shared__ __align__(16) unsigned short W[DO * LD];
  const int node0 = blockIdx.x * gnn::TILE;
  for (int q = threadIdx.x; q < DO * K; q += 256) {
    const int o = q / K, k = q - o * K;
    W[o * LD + k] = wt[o * K + k];
  }
  build_inputs<DI>(n, row, col, hin, node0, X, LD);
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int nb = 0; nb < DO / 16; nb++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma_block<K>(X + (16 * w) * LD, LD, W + (16 * nb) * LD, LD, acc);
    const int cc = 16 * nb + (l & 15);
    const float bb = bias[cc];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int node = node0 + 16 * w + 4 * (l >> 4) + i;
      if (node >= n) continue;
      float z = acc[i] + bb;
      if (ACT) {
        hout[(size_t)node * DO + cc] = f2bf(tanhf(z));
      } else if (cc < 3) {
        out3[(size_t)node * 3 + cc] = z;
      }
    }
  }
}

// Backward of one layer: given g (gradient w.r.t. the layer output, fp32,
// N x DOG where DOG = DO for hidden layers, 3 for the last), writes per-tile
// partials of dW (2DI x DO) and db (DO), and gX split into gself (N x DI) and
// gagg (N x DI, already divided by the node degree).
template <int DI, int DO, bool ACT, int DOG>
__global__ __launch_bounds__(256) void k_bwd(int n, const int* __restrict__ row,
                                             const int* __restrict__ col,
                                             const unsigned short* __restrict__ hin,
                                             const unsigned short* __restrict__ hout,
                                             const float* __restrict__ g,
                                             const unsigned short* __restrict__ wn,
                                             float* __restrict__ part, float* __restrict__ gself,
                                             float* __restrict__ gagg, int want_gx) {
  constexpr int K = 2 * DI;
  constexpr int LDX = K + 4;
  constexpr int LDT = gnn::TILE + 4;
  constexpr int LDG = DO + 4;
  __shared__ __align__(16) unsigned short X[gnn::TILE * LDX];   // inputs, row-major
  __shared__ __align__(16) unsigned short Xt[K * LDT];          // inputs transposed
  __shared__ __align__(16) unsigned short GZ[gnn::TILE * LDG];  // gz row-major
  __shared__ __align__(16) unsigned short GZt[DO * LDT];        // gz transposed
  __shared__ __align__(16) unsigned short Wn[K * LDG];          // W [K][DO]
  __shared__ float dbs[4][DO];
  const int node0 = blockIdx.x * gnn::TILE;
  for (int q = threadIdx.x; q < K * DO; q += 256) {
    const int k = q / DO, o = q - k * DO;
    Wn[k * LDG + o] = wn[k * DO + o];
  }
  build_inputs<DI>(n, row, col, hin, node0, X, LDX);
  // gz = g * act'(z)
  for (int q = threadIdx.x; q < gnn::TILE * DO; q += 256) {
    const int r = q / DO, o = q - r * DO;
    const int node = node0 + r;
    float v = 0.0f;
    if (node < n && o < DOG) {
      v = g[(size_t)node * DOG + o];
      if (ACT) {
        const float hv = bf2f(hout[(size_t)node * DO + o]);
        v *= 1.0f - hv * hv;
      }
    }
    const unsigned short bv = f2bf(v);
    GZ[r * LDG + o] = bv;
    GZt[o * LDT + r] = bv;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < gnn::TILE * K; q += 256) {
    const int r = q / K, k = q - r * K;
    Xt[k * LDT + r] = X[r * LDX + k];
  }
  // db partial (fp32 sums of the bf16-rounded gz, like the MFMA operands)
  {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int o = l; o < DO; o += 64) {
      float s = 0.0f;
      for (int r = 16 * w; r < 16 * w + 16; r++) s += bf2f(GZ[r * LDG + o]);
      dbs[w][o] = s;
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  float* pw = part + (size_t)blockIdx.x * (K * DO + DO);
  // dW = X^T gz : (K x 64) (64 x DO); blocks (K/16) x (DO/16) over 4 waves
  constexpr int MB = K / 16, NB = DO / 16;
  for (int blk = w; blk < MB * NB; blk += 4) {
    const int mb = blk / NB, nb = blk - mb * NB;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma_block<gnn::TILE>(Xt + (16 * mb) * LDT, LDT, GZt + (16 * nb) * LDT, LDT, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = 16 * mb + 4 * (l >> 4) + i, o = 16 * nb + (l & 15);
      pw[k * DO + o] = acc[i];
    }
  }
  if (threadIdx.x < DO) pw[K * DO + threadIdx.x] = dbs[0][threadIdx.x] + dbs[1][threadIdx.x] +
                                                  dbs[2][threadIdx.x] + dbs[3][threadIdx.x];
  if (!want_gx) return;
  // gX = gz W^T : (64 x DO) (DO x K); Bt = W rows (K x DO)
#pragma unroll
  for (int kb = 0; kb < K / 16; kb++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma_block<DO>(GZ + (16 * w) * LDG, LDG, Wn + (16 * kb) * LDG, LDG, acc);
    const int k = 16 * kb + (l & 15);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int node = node0 + 16 * w + 4 * (l >> 4) + i;
      if (node >= n) continue;
      if (k < DI) {
        gself[(size_t)node * DI + k] = acc[i];
      } else {
        const int deg = row[node + 1] - row[node];
        gagg[(size_t)node * DI + (k - DI)] = deg > 0 ? acc[i] / (float)deg : 0.0f;
      }
    }
  }
}

// g_in_j = gself_j + sum_{i in N(j)} gagg_i
template <int DI>
__global__ __launch_bounds__(256) void k_bwd_gather(int n, const int* __restrict__ row,
                                                    const int* __restrict__ col,
                                                    const float* __restrict__ gself,
                                                    const float* __restrict__ gagg,
                                                    float* __restrict__ gin) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int node = t / DI, c = t - (t / DI) * DI;
  if (node >= n) return;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  const int re = row[node + 1];
  int b = row[node];
  for (; b + 4 <= re; b += 4) {  // four independent loads in flight
    s0 += gagg[(size_t)col[b] * DI + c];
    s1 += gagg[(size_t)col[b + 1] * DI + c];
    s2 += gagg[(size_t)col[b + 2] * DI + c];
    s3 += gagg[(size_t)col[b + 3] * DI + c];
  }
  for (; b < re; b++) s0 += gagg[(size_t)col[b] * DI + c];
  gin[(size_t)node * DI + c] = gself[(size_t)node * DI + c] + ((s0 + s1) + (s2 + s3));
}

// Deterministic sum of per-tile partials into the master gradient layout.
__global__ __launch_bounds__(256) void k_reduce_layer(int ntiles, const float* __restrict__ part,
                                                      int di, int dip, int dout, int doutp,
                                                      float* __restrict__ grad) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int np = 2 * di * dout + dout;
  if (q >= np) return;
  int src;
  if (q < 2 * di * dout) {
    const int half = q / (di * dout);      // 0: self rows, 1: neighbour rows
    const int rem = q - half * di * dout;
    const int k = rem / dout, o = rem - (rem / dout) * dout;
    src = (half * dip + k) * doutp + o;
  } else {
    src = 2 * dip * doutp + (q - 2 * di * dout);
  }
  const int stride = 2 * dip * doutp + doutp;
  float s = 0.0f;
  for (int t = 0; t < ntiles; t++) s += part[(size_t)t * stride + src];
  grad[q] = s;
}

// Packs master FP32 params of a layer into the padded bf16 forward/backward
// weight layouts and the padded bias.
__global__ void k_pack(const float* __restrict__ p, int di, int dip, int dout, int doutp,
                       unsigned short* __restrict__ wt, unsigned short* __restrict__ wn,
                       float* __restrict__ bias) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int K = 2 * dip;
  if (q < K * doutp) {
    const int k = q / doutp, o = q - (q / doutp) * doutp;
    const int half = k / dip, kk = k - half * dip;
    float v = 0.0f;
    if (kk < di && o < dout) v = p[half * di * dout + kk * dout + o];
    const unsigned short b = f2bf(v);
    wn[k * doutp + o] = b;
    wt[o * K + k] = b;
  }
  if (q < doutp) bias[q] = q < dout ? p[2 * di * dout + q] : 0.0f;
}

__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1t,
                       float b2t) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= np) return;
  const float gg = g[q];
  const float mm = 0.9f * m[q] + 0.1f * gg;
  const float vv = 0.999f * v[q] + 0.001f * gg * gg;
  m[q] = mm;
  v[q] = vv;
  p[q] -= lr * (mm / (1.0f - b1t)) / (sqrtf(vv / (1.0f - b2t)) + 1e-8f);
}

// z0 = D^-1 r; partial sums of r.r and z0.z0
__global__ __launch_bounds__(256) void k_feat_pre(int n, const double* __restrict__ dinv,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ z0,
                                                  double* __restrict__ red, int nred) {
  __shared__ double sh[2][4];
  double sr = 0.0, sz = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double* D = dinv + 9 * (size_t)i;
    const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
    const double a = D[0] * r0 + D[1] * r1 + D[2] * r2;
    const double b = D[3] * r0 + D[4] * r1 + D[5] * r2;
    const double c = D[6] * r0 + D[7] * r1 + D[8] * r2;
    z0[3 * i] = a; z0[3 * i + 1] = b; z0[3 * i + 2] = c;
    sr += r0 * r0 + r1 * r1 + r2 * r2;
    sz += a * a + b * b + c * c;
  }
  for (int o = 32; o > 0; o >>= 1) { sr += __shfl_xor(sr, o); sz += __shfl_xor(sz, o); }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = sr; sh[1][w] = sz; }
  __syncthreads();
  if (threadIdx.x == 0) {
    red[blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    red[nred + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

__global__ void k_feat_scales(int nred, int n3, const double* __restrict__ red, float* scales) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int k = 0; k < nred; k++) { a += red[k]; b += red[nred + k]; }
  scales[0] = (float)sqrt(a / n3);
  scales[1] = (float)sqrt(b / n3);
}

__global__ __launch_bounds__(256) void k_features(int n, const int* __restrict__ row,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ z0,
                                                  const signed char* __restrict__ fixed,
                                                  const float* __restrict__ scales,
                                                  unsigned short* __restrict__ x) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float ir = scales[0] > 0.0f ? 1.0f / scales[0] : 0.0f;
  const float iz = scales[1] > 0.0f ? 1.0f / scales[1] : 0.0f;
  unsigned short* xi = x + (size_t)i * 16;
  for (int c = 0; c < 3; c++) {
    xi[c] = f2bf((float)r[3 * i + c] * ir);
    xi[3 + c] = f2bf((float)z0[3 * i + c] * iz);
  }
  xi[6] = f2bf((fixed[3 * i] || fixed[3 * i + 1] || fixed[3 * i + 2]) ? 1.0f : 0.0f);
  xi[7] = f2bf((float)(row[i + 1] - row[i]) * (1.0f / 128.0f));
  for (int c = 8; c < 16; c++) xi[c] = 0;
}

__global__ void k_output(int n3, const double* __restrict__ z0, const float* __restrict__ out,
                         const signed char* __restrict__ fixed, const float* __restrict__ scales,
                         double* __restrict__ z) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n3) return;
  z[q] = fixed[q] ? 0.0 : z0[q] + (double)scales[1] * (double)out[q];
}

// target t = (e - z0) / s_z and loss gradient g = 2 (o - t) / sum(t^2)
// Training loss: squared energy-norm error of the preconditioned step,
//   L = (z - e)^T K (z - e) / (e^T K e),  z = D^-1 r + s_z o,  e = K^-1 r,
// so dL/do = 2 s_z K (z - e) / (e^T K e) on free DOFs (K symmetric).
__global__ void k_step_error(int n3, const double* __restrict__ e, const double* __restrict__ z0,
                             const float* __restrict__ o, const signed char* __restrict__ fixed,
                             const float* __restrict__ scales, double* __restrict__ dz) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n3) return;
  dz[q] = fixed[q] ? 0.0 : z0[q] + (double)scales[1] * (double)o[q] - e[q];
}

__global__ void k_masked_copy(int n3, const double* __restrict__ src,
                              const signed char* __restrict__ fixed, double* __restrict__ dst) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < n3) dst[q] = fixed[q] ? 0.0 : src[q];
}

__global__ __launch_bounds__(256) void k_loss_grad(int n3, const double* __restrict__ dz,
                                                   const double* __restrict__ kdz,
                                                   const signed char* __restrict__ fixed,
                                                   float* __restrict__ g, double* red) {
  __shared__ double sh[4];
  double se = 0.0;
  for (int q = blockIdx.x * 256 + threadIdx.x; q < n3; q += gridDim.x * 256) {
    const double k = fixed[q] ? 0.0 : kdz[q];
    se += dz[q] * k;
    g[q] = (float)k;  // scaled below once the loss normaliser is known
  }
  for (int k = 32; k > 0; k >>= 1) se += __shfl_xor(se, k);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = se;
  __syncthreads();
  if (threadIdx.x == 0) red[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// scales[2] = relative energy-norm error sqrt(L); scales[3] = 2 s_z / e^T K e
__global__ void k_loss_final(int nred, const double* __restrict__ red, double eke, float* scales) {
  if (threadIdx.x != 0) return;
  double a = 0.0;
  for (int k = 0; k < nred; k++) a += red[k];
  const double inv = eke > 0.0 ? 1.0 / eke : 0.0;
  scales[2] = (float)std::sqrt(std::max(a * inv, 0.0));
  scales[3] = (float)(2.0 * (double)scales[1] * inv);
}

__global__ void k_scale_grad(int n3, float* __restrict__ g, const float* __restrict__ scales) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < n3) g[q] *= scales[3];
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace gnnh {

using namespace gnn;

static int ensure(Ctx* c, unsigned seed) {
  if (c->gnn && c->gnn->n == c->N) return PDG_OK;
  c->free_gnn();
  GnnState* G = new GnnState();
  c->gnn_epoch++;
  c->gnn = G;
  const int N = c->N;
  G->n = N;
  G->ntiles = (N + TILE - 1) / TILE;
  G->nred = std::min(512, std::max(1, (N + 255) / 256));
  auto A = [&](void** p, size_t bytes) { return c->hip(hipMalloc(p, std::max<size_t>(bytes, 64)), "gnn alloc"); };
  const int np = total_params();
  int rc = 0;
  if (!rc) rc = A((void**)&G->params, 4 * (size_t)np);
  if (!rc) rc = A((void**)&G->grads, 4 * (size_t)np);
  if (!rc) rc = A((void**)&G->adam_m, 4 * (size_t)np);
  if (!rc) rc = A((void**)&G->adam_v, 4 * (size_t)np);
  for (int l = 0; l < NL && !rc; l++) {
    rc = A((void**)&G->wt[l], 2 * (size_t)DOUTP[l] * 2 * DINP[l]);
    if (!rc) rc = A((void**)&G->wn[l], 2 * (size_t)DOUTP[l] * 2 * DINP[l]);
    if (!rc) rc = A((void**)&G->bias[l], 4 * (size_t)DOUTP[l]);
  }
  if (!rc) rc = A((void**)&G->h[0], 2 * (size_t)N * 16);
  if (!rc) rc = A((void**)&G->h[1], 2 * (size_t)N * 32);
  if (!rc) rc = A((void**)&G->h[2], 2 * (size_t)N * 32);
  if (!rc) rc = A((void**)&G->out, 4 * (size_t)N * 3);
  if (!rc) rc = A((void**)&G->gbuf[0], 4 * (size_t)N * 32);
  if (!rc) rc = A((void**)&G->gbuf[1], 4 * (size_t)N * 32);
  if (!rc) rc = A((void**)&G->gagg, 4 * (size_t)N * 32);
  if (!rc) rc = A((void**)&G->part, 4 * (size_t)G->ntiles * (2 * 32 * 32 + 32));
  if (!rc) rc = A((void**)&G->scales, 64);
  if (!rc) rc = A((void**)&G->z0, 24 * (size_t)N);
  if (!rc) rc = A((void**)&G->dz, 24 * (size_t)N);
  if (!rc) rc = A((void**)&G->kdz, 24 * (size_t)N);
  if (!rc) rc = A((void**)&G->red, 16 * (size_t)G->nred);
  if (rc) return rc;
  // Glorot-style initial weights; the last layer starts at zero so the
  // untrained preconditioner equals block Jacobi.
  std::mt19937 gen(seed);
  std::vector<float> p(np, 0.0f);
  for (int l = 0; l < NL; l++) {
    const int o = param_off(l);
    const float sd = std::sqrt(2.0f / (float)(2 * DIN[l] + DOUT[l]));
    std::normal_distribution<float> nd(0.0f, sd);
    for (int q = 0; q < 2 * DIN[l] * DOUT[l]; q++) p[o + q] = l == NL - 1 ? 0.0f : nd(gen);
  }
  rc = c->hip(hipMemcpyAsync(G->params, p.data(), 4 * (size_t)np, hipMemcpyHostToDevice, c->stream), "h2d");
  if (!rc) rc = c->hip(hipMemsetAsync(G->adam_m, 0, 4 * (size_t)np, c->stream), "zero");
  if (!rc) rc = c->hip(hipMemsetAsync(G->adam_v, 0, 4 * (size_t)np, c->stream), "zero");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  G->step = 0;
  return rc;
}

static int pack(Ctx* c) {
  GnnState* G = c->gnn;
  for (int l = 0; l < NL; l++) {
    const long long cnt = (long long)2 * DINP[l] * DOUTP[l];
    hipLaunchKernelGGL(k_pack, dim3(nblk(cnt)), dim3(256), 0, c->stream, G->params + param_off(l),
                       DIN[l], DINP[l], DOUT[l], DOUTP[l], G->wt[l], G->wn[l], G->bias[l]);
  }
  return c->hip(hipGetLastError(), "gnn pack");
}

// features already in G->h[0]; runs the three layers into G->out
static int forward(Ctx* c) {
  GnnState* G = c->gnn;
  const int N = c->N;
  const dim3 grid(G->ntiles), blk(256);
  hipLaunchKernelGGL((k_fwd<16, 32, true>), grid, blk, 0, c->stream, N, c->d_row, c->d_col, G->h[0], G->wt[0], G->bias[0], G->h[1], nullptr);
  hipLaunchKernelGGL((k_fwd<32, 32, true>), grid, blk, 0, c->stream, N, c->d_row, c->d_col, G->h[1], G->wt[1], G->bias[1], G->h[2], nullptr);
  hipLaunchKernelGGL((k_fwd<32, 16, false>), grid, blk, 0, c->stream, N, c->d_row, c->d_col, G->h[2], G->wt[2], G->bias[2], nullptr, G->out);
  return c->hip(hipGetLastError(), "gnn forward");
}

// gradient w.r.t. the network output in G->gbuf[1] (N x 3) -> G->grads
static int backward(Ctx* c) {
  GnnState* G = c->gnn;
  const int N = c->N;
  const dim3 grid(G->ntiles), blk(256);
  float* g_out = G->gbuf[1];
  float* g_next = G->gbuf[0];
  hipStream_t s = c->stream;
  // layer 3 (identity)
  hipLaunchKernelGGL((k_bwd<32, 16, false, 3>), grid, blk, 0, s, N, c->d_row, c->d_col, G->h[2], nullptr, g_out, G->wn[2], G->part, g_next, G->gagg, 1);
  hipLaunchKernelGGL(k_reduce_layer, dim3(nblk(layer_params(2))), blk, 0, s, G->ntiles, G->part, DIN[2], DINP[2], DOUT[2], DOUTP[2], G->grads + param_off(2));
  hipLaunchKernelGGL((k_bwd_gather<32>), dim3(nblk((long long)N * 32)), blk, 0, s, N, c->d_row, c->d_col, g_next, G->gagg, g_out);
  // layer 2
  hipLaunchKernelGGL((k_bwd<32, 32, true, 32>), grid, blk, 0, s, N, c->d_row, c->d_col, G->h[1], G->h[2], g_out, G->wn[1], G->part, g_next, G->gagg, 1);
  hipLaunchKernelGGL(k_reduce_layer, dim3(nblk(layer_params(1))), blk, 0, s, G->ntiles, G->part, DIN[1], DINP[1], DOUT[1], DOUTP[1], G->grads + param_off(1));
  hipLaunchKernelGGL((k_bwd_gather<32>), dim3(nblk((long long)N * 32)), blk, 0, s, N, c->d_row, c->d_col, g_next, G->gagg, g_out);
  // layer 1 (no gradient w.r.t. the features)
  hipLaunchKernelGGL((k_bwd<16, 32, true, 32>), grid, blk, 0, s, N, c->d_row, c->d_col, G->h[0], G->h[1], g_out, G->wn[0], G->part, g_next, G->gagg, 0);
  hipLaunchKernelGGL(k_reduce_layer, dim3(nblk(layer_params(0))), blk, 0, s, G->ntiles, G->part, DIN[0], DINP[0], DOUT[0], DOUTP[0], G->grads + param_off(0));
  return c->hip(hipGetLastError(), "gnn backward");
}

// features from r (device, internal order) using the block-Jacobi inverse in c->qs->dinv
static int features(Ctx* c, const double* r) {
  GnnState* G = c->gnn;
  const int N = c->N;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_feat_pre, dim3(G->nred), dim3(256), 0, s, N, c->qs_dinv(), r, G->z0, G->red, G->nred);
  hipLaunchKernelGGL(k_feat_scales, dim3(1), dim3(64), 0, s, G->nred, 3 * N, G->red, G->scales);
  hipLaunchKernelGGL(k_features, dim3(nblk(N)), dim3(256), 0, s, N, c->d_row, r, G->z0, c->qs_fixed(), G->scales, G->h[0]);
  return c->hip(hipGetLastError(), "gnn features");
}

}  // namespace gnnh

void Ctx::free_gnn() {
  if (!gnn) return;
  GnnState* G = gnn;
  void* ptrs[] = {G->params, G->grads, G->adam_m, G->adam_v, G->h[0], G->h[1], G->h[2], G->out,
                  G->gbuf[0], G->gbuf[1], G->gagg, G->part, G->scales, G->z0, G->red, G->dz, G->kdz};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  for (int l = 0; l < gnn::NL; l++) {
    if (G->wt[l]) (void)hipFree(G->wt[l]);
    if (G->wn[l]) (void)hipFree(G->wn[l]);
    if (G->bias[l]) (void)hipFree(G->bias[l]);
  }
  delete G;
  gnn = nullptr;
}

int gnn_apply(Ctx* c, const double* r, double* z) {
  if (!c->gnn || !c->gnn->trained) return c->fail(PDG_E_ARG, "learned preconditioner not trained");
  GnnState* G = c->gnn;
  int rc = gnnh::features(c, r);
  if (!rc) rc = gnnh::forward(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_output, dim3(nblk(3LL * c->N)), dim3(256), 0, c->stream, 3 * c->N, G->z0, G->out, c->qs_fixed(), G->scales, z);
  return c->hip(hipGetLastError(), "gnn output");
}

}  // namespace pdg

using namespace pdg;

extern "C" {

int pdg_gnn_init(pdg_ctx* h, unsigned seed) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  c->free_gnn();
  int rc = gnnh::ensure(c, seed);
  if (!rc) rc = gnnh::pack(c);
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int pdg_gnn_num_params(void) { return gnn::total_params(); }

int pdg_gnn_get_params(pdg_ctx* h, float* p) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->gnn || !p) return PDG_E_ARG;
  int rc = c->hip(hipMemcpyAsync(p, c->gnn->params, 4 * (size_t)gnn::total_params(), hipMemcpyDeviceToHost, c->stream), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int pdg_gnn_set_params(pdg_ctx* h, const float* p) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !p) return PDG_E_ARG;
  int rc = gnnh::ensure(c, 1);
  if (!rc) rc = c->hip(hipMemcpyAsync(c->gnn->params, p, 4 * (size_t)gnn::total_params(), hipMemcpyHostToDevice, c->stream), "h2d");
  if (!rc) rc = gnnh::pack(c);
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (!rc) c->gnn->trained = true;
  return rc;
}

// Forward on given node features (N x 8, reference order) and the gradient of
// sum(gout * out) w.r.t. the parameters (for checks against a CPU reference).
int pdg_gnn_forward_backward(pdg_ctx* h, const float* feat, const float* gout, float* out,
                             float* dparams) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->gnn || !feat) return PDG_E_ARG;
  GnnState* G = c->gnn;
  const int N = c->N;
  std::vector<unsigned short> x((size_t)N * 16, 0);
  auto tobf = [](float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
  };
  for (int k = 0; k < N; k++)
    for (int q = 0; q < 8; q++) x[(size_t)k * 16 + q] = tobf(feat[(size_t)c->perm[k] * 8 + q]);
  int rc = c->hip(hipMemcpyAsync(G->h[0], x.data(), 2 * x.size(), hipMemcpyHostToDevice, c->stream), "h2d");
  if (!rc) rc = gnnh::pack(c);
  if (!rc) rc = gnnh::forward(c);
  if (rc) return rc;
  std::vector<float> o((size_t)N * 3);
  rc = c->hip(hipMemcpyAsync(o.data(), G->out, 4 * o.size(), hipMemcpyDeviceToHost, c->stream), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  if (out)
    for (int k = 0; k < N; k++)
      for (int q = 0; q < 3; q++) out[(size_t)c->perm[k] * 3 + q] = o[(size_t)k * 3 + q];
  if (gout && dparams) {
    std::vector<float> g((size_t)N * 3);
    for (int k = 0; k < N; k++)
      for (int q = 0; q < 3; q++) g[(size_t)k * 3 + q] = gout[(size_t)c->perm[k] * 3 + q];
    rc = c->hip(hipMemcpyAsync(G->gbuf[1], g.data(), 4 * g.size(), hipMemcpyHostToDevice, c->stream), "h2d");
    if (!rc) rc = gnnh::backward(c);
    if (!rc) rc = c->hip(hipMemcpyAsync(dparams, G->grads, 4 * (size_t)gnn::total_params(), hipMemcpyDeviceToHost, c->stream), "d2h");
    if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  }
  return rc;
}

// Trains on (r, e = K^-1 r) pairs, vectors in reference order (nsamples x 3N),
// with the block-Jacobi inverse at u. Loss: ||o - (e - D^-1 r)/s_z||^2 /
// ||(e - D^-1 r)/s_z||^2 per sample. Returns the mean loss of the last epoch.
int pdg_gnn_train(pdg_ctx* h, const double* u, int nsamples, const double* r, const double* e,
                  int epochs, float lr, unsigned seed, float* lastLoss) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || nsamples <= 0 || !r || !e || epochs <= 0) return PDG_E_ARG;
  int rc = pdg_qs_prepare(c, u);  // block Jacobi at u
  if (rc) return rc;
  if (!c->gnn) rc = gnnh::ensure(c, seed);
  if (rc) return rc;
  GnnState* G = c->gnn;
  const int N = c->N, n3 = 3 * N;
  hipStream_t s = c->stream;
  // samples resident on the device, internal order
  double *dr = nullptr, *de = nullptr;
  rc = c->hip(hipMalloc((void**)&dr, 8 * (size_t)n3 * nsamples), "alloc");
  if (!rc) rc = c->hip(hipMalloc((void**)&de, 8 * (size_t)n3 * nsamples), "alloc");
  if (rc) { if (dr) (void)hipFree(dr); return rc; }
  {
    std::vector<double> t((size_t)n3);
    for (int smp = 0; smp < nsamples && !rc; smp++) {
      for (int pass = 0; pass < 2 && !rc; pass++) {
        const double* src = (pass == 0 ? r : e) + (size_t)smp * n3;
        for (int k = 0; k < N; k++)
          for (int q = 0; q < 3; q++) t[3 * (size_t)k + q] = src[3 * (size_t)c->perm[k] + q];
        rc = c->hip(hipMemcpy((pass == 0 ? dr : de) + (size_t)smp * n3, t.data(), 8 * (size_t)n3, hipMemcpyHostToDevice), "h2d");
      }
    }
  }
  // energy norms e^T K e of the targets (loss normalisers)
  std::vector<double> eke(nsamples, 0.0);
  for (int smp = 0; smp < nsamples && !rc; smp++) {
    const double* es = de + (size_t)smp * n3;
    hipLaunchKernelGGL(k_masked_copy, dim3(nblk(n3)), dim3(256), 0, s, n3, es, c->qs_fixed(), G->dz);
    rc = qs_tangent_apply(c, G->dz, G->kdz);
    if (!rc) rc = qs_dot(c, G->dz, G->kdz, &eke[smp]);
  }
  std::mt19937 gen(seed);
  std::vector<int> order(nsamples);
  for (int k = 0; k < nsamples; k++) order[k] = k;
  const int np = gnn::total_params();
  double epoch_loss = 0.0;
  for (int ep = 0; ep < epochs && !rc; ep++) {
    std::shuffle(order.begin(), order.end(), gen);
    epoch_loss = 0.0;
    for (int q = 0; q < nsamples && !rc; q++) {
      const double* rs = dr + (size_t)order[q] * n3;
      const double* es = de + (size_t)order[q] * n3;
      rc = gnnh::pack(c);
      if (!rc) rc = gnnh::features(c, rs);
      if (!rc) rc = gnnh::forward(c);
      if (rc) break;
      hipLaunchKernelGGL(k_step_error, dim3(nblk(n3)), dim3(256), 0, s, n3, es, G->z0, G->out, c->qs_fixed(), G->scales, G->dz);
      rc = qs_tangent_apply(c, G->dz, G->kdz);
      if (rc) break;
      hipLaunchKernelGGL(k_loss_grad, dim3(G->nred), dim3(256), 0, s, n3, G->dz, G->kdz, c->qs_fixed(), G->gbuf[1], G->red);
      hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, s, G->nred, G->red, eke[order[q]], G->scales);
      hipLaunchKernelGGL(k_scale_grad, dim3(nblk(n3)), dim3(256), 0, s, n3, G->gbuf[1], G->scales);
      rc = gnnh::backward(c);
      if (rc) break;
      G->step++;
      const float b1t = std::pow(0.9f, (float)G->step), b2t = std::pow(0.999f, (float)G->step);
      hipLaunchKernelGGL(k_adam, dim3(nblk(np)), dim3(256), 0, s, np, G->params, G->grads, G->adam_m, G->adam_v, lr, b1t, b2t);
      float sc[4];
      rc = c->hip(hipMemcpyAsync(sc, G->scales, 16, hipMemcpyDeviceToHost, s), "d2h");
      if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
      epoch_loss += sc[2];
    }
    epoch_loss /= nsamples;
  }
  (void)hipFree(dr);
  (void)hipFree(de);
  if (rc) return rc;
  rc = gnnh::pack(c);
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  G->trained = true;
  if (lastLoss) *lastLoss = (float)epoch_loss;
  return rc;
}

}  // extern "C"
