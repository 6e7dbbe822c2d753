// Quasi-static path: FP64 residual, matrix-free tangent apply and 3x3
// block-Jacobi preconditioner of the LPS model, device PCG / flexible CG, and
// a Newton driver over load steps.
//
// Reference: computeQuasiStaticResidual (src/core/Peridigm.cpp:4176-4317):
// r = (f(u) + f_ext) V with the rows of kinematic DOFs zeroed; the Newton
// tangent (Sacado AD over the same templated force kernel,
// src/materials/Peridigm_ElasticMaterial.cpp:258-383); the 3x3 block
// preconditioner of the NOX path (src/core/Peridigm.cpp:1603-1647).
//
// Here the tangent is never assembled. K w = -V (df/du) w is the forward-mode
// directional derivative of the both-ends gather force, evaluated with dual
// numbers in FP64 in the same two bond passes as the force (dilatation first,
// then pair forces); prescribed DOFs get identity rows. The diagonal 3x3
// blocks of K are formed analytically from the same bond quantities.
//
// Dual-number derivatives of the LPS force, for bond i->j with n = Y/|Y|,
// dY = w_j - w_i, de = n.dY:
//   d theta_i = 3/m_i sum_j w (1-d) zeta de V_j
//   T = t_ij + t_ji,  dT = w (1-d) [(dtheta_i b_i + dtheta_j b_j) zeta + (a_i + a_j) de]
//   df_i = sum_j [dT n + T (dY - n (n.dY)) / |Y|] V_j
// with a = 15 mu / m, b = 3K/m - a/3.

#include <algorithm>
#include <cmath>
#include <vector>

#include "pdg_ctx.hpp"
#include "pdg_device.hpp"

namespace pdg {

// Regular coarse grid of the two-level preconditioner: d[k] points along axis
// k from o[k] with spacing H; a position's cell is clamped to the grid.
struct CoarseGrid {
  int d[3] = {2, 2, 2};
  double o[3] = {0.0, 0.0, 0.0};
  double H = 1.0;
  int points() const { return d[0] * d[1] * d[2]; }
};

struct QsState {
  int n = 0;
  double *x = nullptr, *u = nullptr, *w = nullptr, *vol = nullptr, *m = nullptr, *alpha = nullptr,
         *beta = nullptr, *theta = nullptr, *dtheta = nullptr, *g = nullptr, *out = nullptr,
         *dinv = nullptr;
  signed char* fixed = nullptr;  // 3n, 1 = prescribed DOF
  // solver vectors (internal order, 3n)
  double *xs = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *b = nullptr,
         *part = nullptr, *scal = nullptr;
  // flexible CG history: directions and their images (FCG_SLOTS each, 3n)
  double* ph[4] = {};
  double* qh[4] = {};
  int nblk = 0;
  double* hscal = nullptr;  // pinned host copy of the solver scalars
  bool graphs = false;      // replay the CG iteration sequences as HIP graphs
  // captured iteration sequences, per preconditioner, sequence (A, B) and
  // flexible-CG history slot; valid for the x and b vectors and the network
  // buffers they were captured with
  hipGraphExec_t gexec[3][2][4] = {};
  const double* gx = nullptr;
  const double* gb = nullptr;
  unsigned gepoch = 0;
  int fallbacks = 0;  // learned-preconditioner solves handed over to block-Jacobi CG
  std::vector<signed char> hfixed;
  // diagonal of K (3n, from the block-Jacobi pass) and the damped-Newton
  // diagonal shift added to every tangent apply while `damped` is set
  double *diag = nullptr, *dampv = nullptr;
  bool damped = false;
  // temperature change per node (internal order) at dT_time, and a 9n
  // partial-stress buffer
  double* dT = nullptr;
  double dT_time = 0.0;
  unsigned dT_epoch = ~0u;
  double* ps = nullptr;
  // load-stepping driver vectors (3n each, allocated by its first run): the
  // displacement at the load step's start, Newton increment, predictor,
  // line-search base state and residual
  double* rv[5] = {};
  // two-level preconditioner: block Jacobi plus a Galerkin coarse correction
  // on a regular coarse grid over the bounding box (three translations per
  // grid point, weighted by the point's trilinear hat function); built at
  // the first two-level solve after the prescribed DOFs are set or a run
  // starts, and kept
  CoarseGrid cgr;
  int nc = 0;
  bool coarse_ok = false;
  int* cstart = nullptr;  // grid points + 1: each point's nodes in clist
  int* clist = nullptr;   // 8n: nodes of the cells around each grid point
  double* wlist = nullptr;  // 8n: the grid point's weight at each entry of clist
  double* cpart = nullptr;  // 3 kCoarseSplit points: restriction partials
  double* cinv = nullptr;    // nc x nc inverse of P^T K P (dropped functions: zero)
  double* cmat = nullptr;    // nc x nc build buffer
  double* cy = nullptr;      // nc: coarse solution
};

namespace {

using namespace dev;

// lanes per node of the tangent, residual and diagonal kernels: 32 on
// meshes below kQsWideNodes nodes (more waves to hide the gathers' latency),
// 16 above (measured on 24,000- and 200,000-node bars)
constexpr int kQsWideNodes = 100000;
constexpr int kScal = 32;   // solver scalar slots on the device
constexpr int kFcgSlots = 4;  // flexible CG keeps kFcgSlots - 1 previous directions
constexpr int RB = 256; // reduction block

__device__ __forceinline__ double infl64(int kind, double zeta, double horizon) {
  if (kind == PDG_INFLUENCE_ONE) return 1.0;
  if (kind == PDG_INFLUENCE_PARABOLIC_DECAY) {
    if (zeta > horizon) return 0.0;
    double s = zeta / horizon;
    return s < 0.5 ? 1.0 : -4.0 * s * s + 4.0 * s;
  }
  double h2 = horizon * horizon * 0.4 * 0.4;
  return exp(-(zeta * zeta) / h2);
}

template <int G>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

struct QsArgs {
  int n;
  const int* row;
  const int* col;
  const int* mw;
  const unsigned long long* mask;
  const double* x;
  const double* u;
  const double* w;
  const double* vol;
  const double* m;
  const double* alpha;
  const double* beta;
  double* theta;
  double* dtheta;
  double* g;
  const signed char* fixed;
  double* out;
  double* dinv;
  double* diag;
  double horizon;
  int influence;
  int linear;  // linearized LPS (force along the reference bond, linear in u)
  const double* dT;  // temperature change per node (thermal strains), or null
  double alphaT;     // thermal expansion coefficient
  double* ps;        // partial stress, 9 per node (force<3>)
};

// dilatation (+ its directional derivative, + d theta_i / d u_i)
template <int QG, bool DUAL, bool DIAG, bool GEN>
__global__ __launch_bounds__(256) void k_qs_theta(QsArgs a) {
  const int lane = threadIdx.x & 63;
  const int gi = lane / QG, l = lane & (QG - 1);
  const int node = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / QG) + gi;
  const bool valid = node < a.n;
  double th = 0.0, dth = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (valid) {
    const int rs = a.row[node], deg = a.row[node + 1] - rs, w0 = a.mw[node];
    const double* xi = a.x + 3 * node;
    const double* ui = a.u + 3 * node;
    for (int k = l; k < deg; k += QG) {
      if ((a.mask[w0 + (k >> 6)] >> (k & 63)) & 1ull) continue;
      const int j = a.col[rs + k];
      const double X0 = a.x[3 * j] - xi[0], X1 = a.x[3 * j + 1] - xi[1], X2 = a.x[3 * j + 2] - xi[2];
      const double zeta = sqrt(X0 * X0 + X1 * X1 + X2 * X2);
      if (GEN && a.linear) {
        // theta_i = 3/m_i sum w V_j xi.(u_j - u_i)
        // (src/materials/Peridigm_LinearLPSPVMaterial.cpp:229-289)
        const double c = infl64(a.influence, zeta, a.horizon) * a.vol[j];
        th += c * (X0 * (a.u[3 * j] - ui[0]) + X1 * (a.u[3 * j + 1] - ui[1]) +
                   X2 * (a.u[3 * j + 2] - ui[2]));
        if (DUAL) {
          const double* wi = a.w + 3 * node;
          dth += c * (X0 * (a.w[3 * j] - wi[0]) + X1 * (a.w[3 * j + 1] - wi[1]) +
                      X2 * (a.w[3 * j + 2] - wi[2]));
        }
        if (DIAG) { g0 -= c * X0; g1 -= c * X1; g2 -= c * X2; }
        continue;
      }
      const double Y0 = X0 + (a.u[3 * j] - ui[0]);
      const double Y1 = X1 + (a.u[3 * j + 1] - ui[1]);
      const double Y2 = X2 + (a.u[3 * j + 2] - ui[2]);
      const double yl = sqrt(Y0 * Y0 + Y1 * Y1 + Y2 * Y2);
      const double iy = 1.0 / yl;
      double e = yl - zeta;
      if (GEN && a.dT) e -= a.alphaT * a.dT[node] * zeta;  // elastic.cxx thermal strain
      const double c = infl64(a.influence, zeta, a.horizon) * zeta * a.vol[j];
      th += c * e;
      const double n0 = Y0 * iy, n1 = Y1 * iy, n2 = Y2 * iy;
      if (DUAL) {
        const double* wi = a.w + 3 * node;
        const double de = n0 * (a.w[3 * j] - wi[0]) + n1 * (a.w[3 * j + 1] - wi[1]) +
                          n2 * (a.w[3 * j + 2] - wi[2]);
        dth += c * de;
      }
      if (DIAG) { g0 -= c * n0; g1 -= c * n1; g2 -= c * n2; }
    }
  }
  th = gsum<QG>(th);
  if (DUAL) dth = gsum<QG>(dth);
  if (DIAG) { g0 = gsum<QG>(g0); g1 = gsum<QG>(g1); g2 = gsum<QG>(g2); }
  if (valid && l == 0) {
    const double s = a.m[node] > 0.0 ? 3.0 / a.m[node] : 0.0;
    a.theta[node] = s * th;
    if (DUAL) a.dtheta[node] = s * dth;
    if (DIAG) { a.g[3 * node] = s * g0; a.g[3 * node + 1] = s * g1; a.g[3 * node + 2] = s * g2; }
  }
}

// MODE 0: residual r = f V (0 on fixed DOFs); MODE 1: K w = -V df (w on fixed);
// MODE 2: f V on every DOF (the reactions are its kinematic entries);
// MODE 3: the partial stress sum_j f_ij (x_j - x_i)^T V_j of the node's own
// bonds (src/materials/elastic.cxx:132-142), 9 per node into a.ps
// at least 5 waves per SIMD (<= 96 VGPRs; the compiler's own choice, 100,
// allowed 4), with no spills
template <int QG, int MODE, bool GEN>
__global__ __launch_bounds__(256, 5) void k_qs_force(QsArgs a) {
  const int lane = threadIdx.x & 63;
  const int gi = lane / QG, l = lane & (QG - 1);
  const int node = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / QG) + gi;
  const bool valid = node < a.n;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  double P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (valid) {
    const int rs = a.row[node], deg = a.row[node + 1] - rs, w0 = a.mw[node];
    const double* xi = a.x + 3 * node;
    const double* ui = a.u + 3 * node;
    const double ai = a.alpha[node], bi = a.beta[node], thi = a.theta[node];
    const double dthi = MODE == 1 ? a.dtheta[node] : 0.0;
    for (int k = l; k < deg; k += QG) {
      if ((a.mask[w0 + (k >> 6)] >> (k & 63)) & 1ull) continue;
      const int j = a.col[rs + k];
      const double X0 = a.x[3 * j] - xi[0], X1 = a.x[3 * j + 1] - xi[1], X2 = a.x[3 * j + 2] - xi[2];
      const double vj = a.vol[j];
      if (GEN && a.linear) {
        // both ends of bond i-j: w V_j [(theta_i b_i + theta_j b_j)
        //   + (a_i + a_j) xi.(u_j - u_i) / |xi|^2] xi
        // (src/materials/Peridigm_LinearLPSPVMaterial.cpp:296-394)
        const double z2 = X0 * X0 + X1 * X1 + X2 * X2;
        const double om = infl64(a.influence, sqrt(z2), a.horizon);
        const double T2 = om * (ai + a.alpha[j]) / z2;
        double s;
        if (MODE != 1) {
          const double pr = X0 * (a.u[3 * j] - ui[0]) + X1 * (a.u[3 * j + 1] - ui[1]) +
                            X2 * (a.u[3 * j + 2] - ui[2]);
          s = om * (thi * bi + a.theta[j] * a.beta[j]) + T2 * pr;
        } else {
          const double* wi = a.w + 3 * node;
          const double pr = X0 * (a.w[3 * j] - wi[0]) + X1 * (a.w[3 * j + 1] - wi[1]) +
                            X2 * (a.w[3 * j + 2] - wi[2]);
          s = om * (dthi * bi + a.dtheta[j] * a.beta[j]) + T2 * pr;
        }
        f0 += s * X0 * vj; f1 += s * X1 * vj; f2 += s * X2 * vj;
        continue;
      }
      const double Y0 = X0 + (a.u[3 * j] - ui[0]);
      const double Y1 = X1 + (a.u[3 * j + 1] - ui[1]);
      const double Y2 = X2 + (a.u[3 * j + 2] - ui[2]);
      const double zeta = sqrt(X0 * X0 + X1 * X1 + X2 * X2);
      const double yl = sqrt(Y0 * Y0 + Y1 * Y1 + Y2 * Y2);
      const double iy = 1.0 / yl;
      const double e = yl - zeta;
      // thermal strain of each end's bond: e_ij = e - alphaT dT_i zeta
      const double ei = (GEN && a.dT) ? e - a.alphaT * a.dT[node] * zeta : e;
      const double ej = (GEN && a.dT) ? e - a.alphaT * a.dT[j] * zeta : e;
      const double om = infl64(a.influence, zeta, a.horizon);
      const double T = om * ((thi * bi + a.theta[j] * a.beta[j]) * zeta + ai * ei + a.alpha[j] * ej);
      const double n0 = Y0 * iy, n1 = Y1 * iy, n2 = Y2 * iy;
      if (MODE == 3) {
        // the node's own half t_ij only
        const double ti = om * (thi * bi * zeta + ai * ei) * vj;
        const double nn[3] = {n0, n1, n2}, xx[3] = {X0, X1, X2};
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) P[3 * r + c] += ti * nn[r] * xx[c];
      } else if (MODE != 1) {
        f0 += T * n0 * vj; f1 += T * n1 * vj; f2 += T * n2 * vj;
      } else {
        const double* wi = a.w + 3 * node;
        const double d0 = a.w[3 * j] - wi[0], d1 = a.w[3 * j + 1] - wi[1], d2 = a.w[3 * j + 2] - wi[2];
        const double de = n0 * d0 + n1 * d1 + n2 * d2;
        const double dT = om * ((dthi * bi + a.dtheta[j] * a.beta[j]) * zeta + (ai + a.alpha[j]) * de);
        const double s = T * iy;
        f0 += (dT * n0 + s * (d0 - n0 * de)) * vj;
        f1 += (dT * n1 + s * (d1 - n1 * de)) * vj;
        f2 += (dT * n2 + s * (d2 - n2 * de)) * vj;
      }
    }
  }
  if (MODE == 3) {
#pragma unroll
    for (int t = 0; t < 9; t++) P[t] = gsum<QG>(P[t]);
    if (valid && l == 0)
      for (int t = 0; t < 9; t++) a.ps[9 * node + t] = P[t];
    return;
  }
  f0 = gsum<QG>(f0); f1 = gsum<QG>(f1); f2 = gsum<QG>(f2);
  if (!valid || l >= 3) return;
  const int q = 3 * node + l;
  const double f = l == 0 ? f0 : (l == 1 ? f1 : f2);
  const double vi = a.vol[node];
  if (MODE == 0) a.out[q] = a.fixed[q] ? 0.0 : f * vi;
  else if (MODE == 2) a.out[q] = f * vi;
  else a.out[q] = a.fixed[q] ? a.w[q] : -f * vi;
}

// Inverse of the 3x3 diagonal block of K at each node.
template <int QG, bool GEN>
__global__ __launch_bounds__(256) void k_qs_diag(QsArgs a) {
  const int lane = threadIdx.x & 63;
  const int gi = lane / QG, l = lane & (QG - 1);
  const int node = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / QG) + gi;
  const bool valid = node < a.n;
  double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (valid) {
    const int rs = a.row[node], deg = a.row[node + 1] - rs, w0 = a.mw[node];
    const double* xi = a.x + 3 * node;
    const double* ui = a.u + 3 * node;
    const double ai = a.alpha[node], bi = a.beta[node], thi = a.theta[node];
    const double gg[3] = {a.g[3 * node], a.g[3 * node + 1], a.g[3 * node + 2]};
    const double vi = a.vol[node];
    for (int k = l; k < deg; k += QG) {
      if ((a.mask[w0 + (k >> 6)] >> (k & 63)) & 1ull) continue;
      const int j = a.col[rs + k];
      const double X0 = a.x[3 * j] - xi[0], X1 = a.x[3 * j + 1] - xi[1], X2 = a.x[3 * j + 2] - xi[2];
      const double vj = a.vol[j];
      const double cj = a.m[j] > 0.0 ? 3.0 / a.m[j] : 0.0;
      if (GEN && a.linear) {
        // d/du_i of s xi with s = w (theta_i b_i + theta_j b_j) + T2 xi.(u_j - u_i):
        // d theta_i/du_i = g_i, d theta_j/du_i = -(3/m_j) w V_i xi
        const double X[3] = {X0, X1, X2};
        const double z2 = X0 * X0 + X1 * X1 + X2 * X2;
        const double om = infl64(a.influence, sqrt(z2), a.horizon);
        const double T2 = om * (ai + a.alpha[j]) / z2;
        const double hj = a.beta[j] * cj * om * vi;
        double gs[3];
        for (int c = 0; c < 3; c++) gs[c] = om * (bi * gg[c] - hj * X[c]) - T2 * X[c];
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) D[3 * r + c] += vj * X[r] * gs[c];
        continue;
      }
      const double Y0 = X0 + (a.u[3 * j] - ui[0]);
      const double Y1 = X1 + (a.u[3 * j + 1] - ui[1]);
      const double Y2 = X2 + (a.u[3 * j + 2] - ui[2]);
      const double zeta = sqrt(X0 * X0 + X1 * X1 + X2 * X2);
      const double yl = sqrt(Y0 * Y0 + Y1 * Y1 + Y2 * Y2);
      const double e = yl - zeta;
      const double ei = (GEN && a.dT) ? e - a.alphaT * a.dT[node] * zeta : e;
      const double ej = (GEN && a.dT) ? e - a.alphaT * a.dT[j] * zeta : e;
      const double om = infl64(a.influence, zeta, a.horizon);
      const double T = om * ((thi * bi + a.theta[j] * a.beta[j]) * zeta + ai * ei + a.alpha[j] * ej);
      const double n[3] = {Y0 / yl, Y1 / yl, Y2 / yl};
      // grad_{u_i}(t_ij + t_ji) = om [b_i zeta g_i - a_i n - b_j zeta^2 om V_i (3/m_j) n - a_j n]
      const double sn = -ai - a.alpha[j] - a.beta[j] * zeta * zeta * om * vi * cj;
      double gt[3];
      for (int c = 0; c < 3; c++) gt[c] = om * (bi * zeta * gg[c] + sn * n[c]);
      const double s = T / yl;
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
          D[3 * r + c] += vj * (n[r] * gt[c] - s * ((r == c ? 1.0 : 0.0) - n[r] * n[c]));
    }
  }
#pragma unroll
  for (int t = 0; t < 9; t++) D[t] = gsum<QG>(D[t]);
  if (!valid || l != 0) return;
  const double vi = a.vol[node];
  double K[9];
  for (int t = 0; t < 9; t++) K[t] = -vi * D[t];
  for (int c = 0; c < 3; c++)
    if (a.fixed[3 * node + c])
      for (int t = 0; t < 3; t++) { K[3 * c + t] = c == t ? 1.0 : 0.0; K[3 * t + c] = c == t ? 1.0 : 0.0; }
  for (int c = 0; c < 3; c++) a.diag[3 * node + c] = K[4 * c];
  // closed-form inverse by cofactors
  const double c00 = K[4] * K[8] - K[5] * K[7], c01 = K[5] * K[6] - K[3] * K[8], c02 = K[3] * K[7] - K[4] * K[6];
  const double det = K[0] * c00 + K[1] * c01 + K[2] * c02;
  double* Di = a.dinv + 9 * node;
  if (!(fabs(det) > 0.0)) {
    for (int t = 0; t < 9; t++) Di[t] = (t % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  const double id = 1.0 / det;
  Di[0] = c00 * id; Di[1] = (K[2] * K[7] - K[1] * K[8]) * id; Di[2] = (K[1] * K[5] - K[2] * K[4]) * id;
  Di[3] = c01 * id; Di[4] = (K[0] * K[8] - K[2] * K[6]) * id; Di[5] = (K[2] * K[3] - K[0] * K[5]) * id;
  Di[6] = c02 * id; Di[7] = (K[1] * K[6] - K[0] * K[7]) * id; Di[8] = (K[0] * K[4] - K[1] * K[3]) * id;
}

// ---- vector kernels (deterministic block reductions, no atomics) ----------
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += sh[k];
  return t;
}

__global__ __launch_bounds__(RB) void k_dot(int n, const double* a, const double* b, double* part) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) s += a[i] * b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RB) void k_final(int nb, const double* part, int stride, double* out) {
  __shared__ double sh[RB / 64];
  for (int c = 0; c < stride; c++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) s += part[c * nb + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = s;
  }
}

// z = Dinv r (block) on 3-vectors of nodes
__global__ __launch_bounds__(RB) void k_block_apply(int n, const double* dinv, const double* r, double* z) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const double* D = dinv + 9 * i;
  const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
  z[3 * i] = D[0] * r0 + D[1] * r1 + D[2] * r2;
  z[3 * i + 1] = D[3] * r0 + D[4] * r1 + D[5] * r2;
  z[3 * i + 2] = D[6] * r0 + D[7] * r1 + D[8] * r2;
}

// z = Dinv r and partial sums of r.z in one pass (grid-stride over nodes, one
// partial per block, like k_dot)
__global__ __launch_bounds__(RB) void k_block_apply_dot(int n, const double* dinv, const double* r,
                                                        double* z, double* part) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
    const double* D = dinv + 9 * i;
    const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
    const double z0 = D[0] * r0 + D[1] * r1 + D[2] * r2;
    const double z1 = D[3] * r0 + D[4] * r1 + D[5] * r2;
    const double z2 = D[6] * r0 + D[7] * r1 + D[8] * r2;
    z[3 * i] = z0;
    z[3 * i + 1] = z1;
    z[3 * i + 2] = z2;
    s += r0 * z0 + r1 * z1 + r2 * z2;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x += al p; r -= al q; partial sums of r.r
// Solver scalars kept on the device (slots of `sc`), so a CG iteration needs
// no host round trip: RZ = r.z, PQ = p.Kp, RR = r.r, RZN = new r.z, ZQ = z.Kp,
// BETA, BRK = 1 after a breakdown (p.Kp <= 0).
enum { SC_RZ = 0, SC_PQ = 1, SC_RR = 2, SC_RZN = 3, SC_ZQ = 4, SC_BETA = 5, SC_BRK = 6 };
// flexible CG: PR = p.r, then per history slot k: PQH + k = p_k.Kp_k, ZQH + k = z.Kp_k
enum { SC_PR = 7, SC_PQH = 8, SC_ZQH = 12 };

// x += al p, r -= al q, partial sums of r.r (into rr_part); al = RZ / PQ (0
// after a breakdown), PQ summed from the partials of k_dot (pq_part)
__global__ __launch_bounds__(RB) void k_cg_step(int n, int nb, double* sc, const double* pq_part,
                                                const double* p, const double* q, double* x,
                                                double* r, double* rr_part) {
  // p.Kp from k_dot's partials, summed by every block in k_final's order (so
  // the same value); block 0 stores it for the breakdown test
  __shared__ double sh[RB / 64];
  __shared__ double pq_s;
  double t = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) t += pq_part[i];
  t = block_sum(t, sh);
  if (threadIdx.x == 0) {
    pq_s = t;
    if (blockIdx.x == 0) sc[SC_PQ] = t;
  }
  __syncthreads();
  const double pq = pq_s;
  // r.z of this iteration is in RZN (from k_cg_dir or the set-up); block 0
  // moves it to RZ for the next direction's beta (no block reads RZ here)
  const double al = pq > 0.0 ? sc[SC_RZN] / pq : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_RZ] = sc[SC_RZN];
  double s = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
    x[i] += al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) rr_part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RB) void k_fcg_step(int n, double* sc, int slot, const double* p,
                                                 const double* q, double* x, double* r,
                                                 double* part) {
  __shared__ double sh[RB / 64];
  const double pq = sc[SC_PQH + slot];
  if (blockIdx.x == 0 && threadIdx.x == 0 && !(pq > 0.0)) sc[SC_BRK] = 1.0;
  const double al = pq > 0.0 ? sc[SC_PR] / pq : 0.0;
  double s = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
    x[i] += al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// New flexible-CG direction: p_new = z - sum_k (ZQH[k] / PQH[k]) p_k over the
// nh history slots listed in `slots` (Notay's truncated FCG).
struct FcgHist {
  const double* p[3];
  int slot[3];
  int nh;
};
__global__ __launch_bounds__(RB) void k_fcg_dir(int n, const double* sc, const double* z,
                                                FcgHist h, double* pnew) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  double v = z[i];
  for (int k = 0; k < h.nh; k++) {
    const double pq = sc[SC_PQH + h.slot[k]];
    const double c = pq > 0.0 ? sc[SC_ZQH + h.slot[k]] / pq : 0.0;
    v -= c * h.p[k][i];
  }
  pnew[i] = v;
}

// p = z + beta p
// p = z + beta p with beta = RZN / RZ (grid-stride, launched with as many
// blocks as there are partials): every block sums the r.z partials
// (rz_part) in the order of a single-block reduction; block 0 stores the sum
// in RZN (read by the next step; no block reads RZN here), BETA, and the
// breakdown flag
__global__ __launch_bounds__(RB) void k_cg_dir(int n, int nb, double* sc, const double* rz_part,
                                               const double* z, double* p) {
  __shared__ double sh[RB / 64];
  __shared__ double beta_s;
  double t = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) t += rz_part[i];
  t = block_sum(t, sh);
  if (threadIdx.x == 0) {
    const double beta = t / sc[SC_RZ];
    beta_s = beta;
    if (blockIdx.x == 0) {
      sc[SC_RZN] = t;
      sc[SC_BETA] = beta;
      if (!(sc[SC_PQ] > 0.0)) sc[SC_BRK] = 1.0;
    }
  }
  __syncthreads();
  const double beta = beta_s;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(RB) void k_axpby(int n, double a, const double* x, double b, double* y) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(RB) void k_sub(int n, const double* a, const double* b, double* c) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) c[i] = a[i] - b[i];
}

// out += s * w (damped-Newton diagonal shift)
__global__ __launch_bounds__(RB) void k_damp_add(int n, const double* sdiag, const double* w,
                                                 double* out) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) out[i] += sdiag[i] * w[i];
}

// s = (f - 1) d + g: the change of the diagonal d scaled by f and shifted by g
__global__ __launch_bounds__(RB) void k_damp_vec(int n, const double* d, double f, double g,
                                                 double* sdiag) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) sdiag[i] = (f * d[i] + g) - d[i];
}

// per-block partial sums of squares and maxima of |v| (optionally only where
// mask != 0): part[b] and part[nb + b]
__global__ __launch_bounds__(RB) void k_norm_part(int n, const double* v, const signed char* mask,
                                                  double* part) {
  __shared__ double sh[RB / 64];
  __shared__ double shm[RB / 64];
  double s = 0.0, m = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
    if (mask && !mask[i]) continue;
    const double x = v[i];
    s += x * x;
    m = fmax(m, fabs(x));
  }
  s = block_sum(s, sh);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < RB / 64; k++) m = fmax(m, shm[k]);
    part[blockIdx.x] = s;
    part[gridDim.x + blockIdx.x] = m;
  }
}

// out[0] = sum of the nb partial sums, out[1] = max of the nb partial maxima
__global__ __launch_bounds__(RB) void k_norm_final(int nb, const double* part, double* out) {
  __shared__ double sh[RB / 64];
  __shared__ double shm[RB / 64];
  double s = 0.0, m = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) {
    s += part[i];
    m = fmax(m, part[nb + i]);
  }
  s = block_sum(s, sh);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < RB / 64; k++) m = fmax(m, shm[k]);
    out[0] = s;
    out[1] = m;
  }
}

// y = a x + b z (z may alias y)
__global__ __launch_bounds__(RB) void k_lin2(int n, double a, const double* x, double b,
                                             const double* z, double* y) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * z[i];
}

// v = 0 on the kinematic DOFs
// Two-level preconditioner kernels. Coarse function 3g + k is the trilinear
// hat function of grid point g times the unit vector along axis k; it is
// zero on prescribed DOFs.
constexpr int kCoarsePointsMax = 160;  // grid points at most (480 coarse functions)
// grid points allowed for n nodes: 64, rising to kCoarsePointsMax at 80,000
// nodes (the probed set-up costs at most 3 x 125 tangent applies)
inline int coarse_points_cap(int n) { return std::max(64, std::min(kCoarsePointsMax, n / 500)); }

// Cell of a position (lower corner, clamped so the cell lies inside the
// grid) and the position's local coordinates in it.
__device__ __forceinline__ void coarse_cell(const CoarseGrid& G, const double* p, int c[3],
                                            double t[3]) {
  for (int k = 0; k < 3; k++) {
    const double f = (p[k] - G.o[k]) / G.H;
    int ci = (int)floor(f);
    ci = ci < 0 ? 0 : (ci > G.d[k] - 2 ? G.d[k] - 2 : ci);
    t[k] = f - ci;
    c[k] = ci;
  }
}

// Weight of corner q (bits: x, y, z offsets) of the cell at local t
__device__ __forceinline__ double corner_w(int q, const double t[3]) {
  return ((q & 1) ? t[0] : 1.0 - t[0]) * ((q & 2) ? t[1] : 1.0 - t[1]) *
         ((q & 4) ? t[2] : 1.0 - t[2]);
}

__device__ __forceinline__ int corner_id(const CoarseGrid& G, const int c[3], int q) {
  return (c[0] + (q & 1)) + G.d[0] * ((c[1] + ((q >> 1) & 1)) + G.d[1] * (c[2] + ((q >> 2) & 1)));
}

// Weight of grid point g at node i (0 when g is not a corner of i's cell)
__device__ __forceinline__ double point_w(const CoarseGrid& G, const double* p, int g) {
  int c[3];
  double t[3];
  coarse_cell(G, p, c, t);
  double w = 0.0;
  for (int q = 0; q < 8; q++) w += corner_id(G, c, q) == g ? corner_w(q, t) : 0.0;
  return w;
}

// Weights of the restriction lists: wlist[e] = weight of grid point pt at
// node clist[e], for the entries e of pt's list (one block per point, once
// per set-up)
__global__ __launch_bounds__(RB) void k_coarse_weights(CoarseGrid G, const int* __restrict__ start,
                                                       const int* __restrict__ list,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ wlist) {
  const int pt = blockIdx.x;
  for (int e = start[pt] + threadIdx.x; e < start[pt + 1]; e += RB)
    wlist[e] = point_w(G, x + 3 * (size_t)list[e], pt);
}

constexpr int kCoarseSplit = 8;  // blocks per grid point in the restriction

// Restriction partials: block (pt, sp) sums phi_(pt,k) . r over the sp-th
// eighth of pt's list into part[3 (pt kCoarseSplit + sp) + k]
__global__ __launch_bounds__(RB) void k_coarse_partial(const int* __restrict__ start,
                                                       const int* __restrict__ list,
                                                       const double* __restrict__ wlist,
                                                       const signed char* __restrict__ fixed,
                                                       const double* __restrict__ r,
                                                       double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  const int pt = blockIdx.x / kCoarseSplit, sp = blockIdx.x - kCoarseSplit * pt;
  const int e0 = start[pt], len = start[pt + 1] - e0;
  const int b0 = e0 + (int)((long long)len * sp / kCoarseSplit);
  const int b1 = e0 + (int)((long long)len * (sp + 1) / kCoarseSplit);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int e = b0 + threadIdx.x; e < b1; e += RB) {
    const int i = list[e];
    const double w = wlist[e];
    for (int k = 0; k < 3; k++) acc[k] += fixed[3 * i + k] ? 0.0 : w * r[3 * i + k];
  }
  for (int k = 0; k < 3; k++) {
    const double t = block_sum(acc[k], sh);
    if (threadIdx.x == 0) part[3 * (size_t)blockIdx.x + k] = t;
  }
}

// g = P^T r from the partials (fixed order), then, when SOLVE, y = Cinv g:
// every block forms g in LDS and its waves take rows (lanes over a row);
// without SOLVE block 0 writes g to `out` (a column of the coarse matrix at
// set-up)
template <bool SOLVE>
__global__ __launch_bounds__(RB) void k_coarse_finish(int nc, const double* __restrict__ part,
                                                      const double* __restrict__ cinv,
                                                      double* __restrict__ out) {
  __shared__ double gl[3 * kCoarsePointsMax];
  if (!SOLVE && blockIdx.x > 0) return;
  for (int t = threadIdx.x; t < nc; t += RB) {
    const int pt = t / 3, k = t - 3 * pt;
    double v = 0.0;
    for (int sp = 0; sp < kCoarseSplit; sp++) v += part[3 * (size_t)(pt * kCoarseSplit + sp) + k];
    if (SOLVE) gl[t] = v;
    else out[t] = v;
  }
  if (!SOLVE) return;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int row = blockIdx.x * (RB / 64) + (threadIdx.x >> 6); row < nc; row += gridDim.x * (RB / 64)) {
    const double* c = cinv + (size_t)row * nc;
    double v = 0.0;
    for (int k = lane; k < nc; k += 64) v += c[k] * gl[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) out[row] = v;
  }
}

// Probe vector of the coarse set-up: the sum of the coarse functions (g, k0)
// over the grid points g of one colour (gx, gy, gz congruent to cx, cy, cz
// modulo `stride`)
__global__ __launch_bounds__(RB) void k_coarse_probe(int n, CoarseGrid G, int stride, int cx, int cy,
                                                     int cz, int k0, const double* __restrict__ x,
                                                     const signed char* __restrict__ fixed,
                                                     double* __restrict__ v) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  int c[3];
  double t[3];
  coarse_cell(G, x + 3 * (size_t)i, c, t);
  double w = 0.0;
  for (int q = 0; q < 8; q++) {
    const bool in = (c[0] + (q & 1)) % stride == cx && (c[1] + ((q >> 1) & 1)) % stride == cy &&
                    (c[2] + ((q >> 2) & 1)) % stride == cz;
    w += in ? corner_w(q, t) : 0.0;
  }
  for (int k = 0; k < 3; k++) v[3 * i + k] = (k == k0 && !fixed[3 * i + k]) ? w : 0.0;
}

// Linear index of each node's cell (the lists of the restriction are built
// from it on the host, so they follow the device's cell assignment)
__global__ __launch_bounds__(RB) void k_coarse_cells(int n, CoarseGrid G, const double* __restrict__ x,
                                                     int* __restrict__ cell) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  int c[3];
  double t[3];
  coarse_cell(G, x + 3 * (size_t)i, c, t);
  cell[i] = c[0] + G.d[0] * (c[1] + G.d[1] * c[2]);
}

// z = Dinv r + P y (the two-level preconditioner), with the partial sums of
// r.z when DOT (grid-stride over nodes, one partial per block, like k_dot)
template <bool DOT>
__global__ __launch_bounds__(RB) void k_two_level_apply(int n, const double* __restrict__ dinv,
                                                        CoarseGrid G, const double* __restrict__ x,
                                                        const signed char* __restrict__ fixed,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ r,
                                                        double* __restrict__ z,
                                                        double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
    const double* D = dinv + 9 * i;
    const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
    int c[3];
    double t[3];
    coarse_cell(G, x + 3 * (size_t)i, c, t);
    double p[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 8; q++) {
      const double w = corner_w(q, t);
      const double* yq = y + 3 * corner_id(G, c, q);
      for (int k = 0; k < 3; k++) p[k] += w * yq[k];
    }
    const double z0 = D[0] * r0 + D[1] * r1 + D[2] * r2 + (fixed[3 * i] ? 0.0 : p[0]);
    const double z1 = D[3] * r0 + D[4] * r1 + D[5] * r2 + (fixed[3 * i + 1] ? 0.0 : p[1]);
    const double z2 = D[6] * r0 + D[7] * r1 + D[8] * r2 + (fixed[3 * i + 2] ? 0.0 : p[2]);
    z[3 * i] = z0;
    z[3 * i + 1] = z1;
    z[3 * i + 2] = z2;
    if (DOT) s += r0 * z0 + r1 * z1 + r2 * z2;
  }
  if (DOT) {
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(RB) void k_zero_fixed(int n, const signed char* fixed, double* v) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n && fixed[i]) v[i] = 0.0;
}

template <int QG>
inline unsigned qblocks(int n) { return (unsigned)((n + 4 * (64 / QG) - 1) / (4 * (64 / QG))); }

// launches with the lane count chosen for the mesh size, and the general
// kernels (GEN: the linearized model or thermal strains) only when needed, so
// the plain elastic loops carry no branch for them
template <bool DUAL, bool DIAG, bool GEN>
void qs_theta_g(const QsArgs& a, hipStream_t s) {
  if (a.n < kQsWideNodes) hipLaunchKernelGGL((k_qs_theta<32, DUAL, DIAG, GEN>), dim3(qblocks<32>(a.n)), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_qs_theta<16, DUAL, DIAG, GEN>), dim3(qblocks<16>(a.n)), dim3(256), 0, s, a);
}
template <bool DUAL, bool DIAG>
void qs_theta(const QsArgs& a, hipStream_t s) {
  if (a.linear || a.dT) qs_theta_g<DUAL, DIAG, true>(a, s);
  else qs_theta_g<DUAL, DIAG, false>(a, s);
}
template <int MODE, bool GEN>
void qs_force_g(const QsArgs& a, hipStream_t s) {
  if (a.n < kQsWideNodes) hipLaunchKernelGGL((k_qs_force<32, MODE, GEN>), dim3(qblocks<32>(a.n)), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_qs_force<16, MODE, GEN>), dim3(qblocks<16>(a.n)), dim3(256), 0, s, a);
}
template <int MODE>
void qs_force(const QsArgs& a, hipStream_t s) {
  if (a.linear || a.dT) qs_force_g<MODE, true>(a, s);
  else qs_force_g<MODE, false>(a, s);
}
template <bool GEN>
void qs_diag_g(const QsArgs& a, hipStream_t s) {
  if (a.n < kQsWideNodes) hipLaunchKernelGGL((k_qs_diag<32, GEN>), dim3(qblocks<32>(a.n)), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_qs_diag<16, GEN>), dim3(qblocks<16>(a.n)), dim3(256), 0, s, a);
}
inline void qs_diag(const QsArgs& a, hipStream_t s) {
  if (a.linear || a.dT) qs_diag_g<true>(a, s);
  else qs_diag_g<false>(a, s);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace qs {

static QsArgs args(Ctx* c) {
  QsState* S = c->qs;
  QsArgs a{};
  a.n = c->N; a.row = c->d_row; a.col = c->d_col; a.mw = c->d_mw; a.mask = c->d_mask;
  a.x = S->x; a.u = S->u; a.w = S->w; a.vol = S->vol; a.m = S->m; a.alpha = S->alpha;
  a.beta = S->beta; a.theta = S->theta; a.dtheta = S->dtheta; a.g = S->g; a.fixed = S->fixed;
  a.out = S->out; a.dinv = S->dinv; a.diag = S->diag; a.horizon = c->horizon;
  a.influence = c->infl;
  a.linear = c->material == PDG_MATERIAL_LINEAR_LPS ? 1 : 0;
  a.dT = c->thermal() ? S->dT : nullptr;
  a.alphaT = c->alpha_t;
  a.ps = S->ps;
  return a;
}

// The temperature change on the device at the context's temperature time.
static int thermal_sync(Ctx* c) {
  QsState* S = c->qs;
  if (!c->thermal()) return PDG_OK;
  const int N = c->N;
  if (!S->dT) {
    int rc = c->hip(hipMalloc((void**)&S->dT, std::max<size_t>(8 * (size_t)N, 64)), "qs alloc");
    if (rc) return rc;
    S->dT_epoch = ~0u;
  }
  if (S->dT_epoch == c->thermal_epoch && S->dT_time == c->temp_time) return PDG_OK;
  std::vector<double> h(N);
  const double t = c->temp_time;
  for (int k = 0; k < N; k++) {
    const int p = c->perm[k];
    h[k] = (c->dT1.empty() ? 0.0 : c->dT1[p]) * t + c->dT0[p];
  }
  int rc = c->hip(hipMemcpyAsync(S->dT, h.data(), 8 * (size_t)N, hipMemcpyHostToDevice, c->stream), "h2d dT");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (!rc) { S->dT_epoch = c->thermal_epoch; S->dT_time = t; }
  return rc;
}

static int ensure(Ctx* c) {
  if (c->material != PDG_MATERIAL_ELASTIC && c->material != PDG_MATERIAL_LINEAR_LPS)
    return c->fail(PDG_E_UNSUPPORTED,
                   "the quasi-static path implements the elastic and linearized LPS models");
  if (c->qs && c->qs->n == c->N) return thermal_sync(c);
  if (c->qs) c->free_extra();
  QsState* S = new QsState();
  c->qs = S;
  const int N = c->N;
  S->n = N;
  S->nblk = std::min(1024, std::max(1, (3 * N + RB - 1) / RB));
  auto A = [&](void** p, size_t bytes) { return c->hip(hipMalloc(p, std::max<size_t>(bytes, 64)), "qs alloc"); };
  int rc = 0;
  double** vecs3[] = {&S->x,  &S->u, &S->w, &S->g, &S->out, &S->xs,   &S->r,
                      &S->z,  &S->p, &S->q, &S->b, &S->diag, &S->dampv};
  for (double** v : vecs3) if (!rc) rc = A((void**)v, 24 * (size_t)N);
  double** vecs1[] = {&S->vol, &S->m, &S->alpha, &S->beta, &S->theta, &S->dtheta};
  for (double** v : vecs1) if (!rc) rc = A((void**)v, 8 * (size_t)N);
  if (!rc) rc = A((void**)&S->dinv, 72 * (size_t)N);
  if (!rc) rc = A((void**)&S->part, 8 * (size_t)S->nblk * 2);
  if (!rc) rc = A((void**)&S->scal, 8 * kScal);
  if (!rc) rc = c->hip(hipHostMalloc((void**)&S->hscal, 8 * kScal, hipHostMallocDefault), "pinned alloc");
  for (int k = 0; k < 4 && !rc; k++) {
    rc = A((void**)&S->ph[k], 24 * (size_t)N);
    if (!rc) rc = A((void**)&S->qh[k], 24 * (size_t)N);
  }
  if (!rc) rc = A((void**)&S->fixed, 3 * (size_t)N);
  if (rc) return rc;
  // constants in internal order
  std::vector<double> hx(3 * (size_t)N), hv(N), hm(N), ha(N), hb(N);
  for (int k = 0; k < N; k++) {
    const int p = c->perm[k];
    for (int d = 0; d < 3; d++) hx[3 * (size_t)k + d] = c->x[3 * (size_t)p + d];
    hv[k] = c->vol[p];
    hm[k] = c->m[p];
    if (hm[k] > 0.0) {
      ha[k] = 15.0 * c->mu / hm[k];
      hb[k] = 3.0 * c->K / hm[k] - ha[k] / 3.0;
    } else {
      ha[k] = hb[k] = 0.0;
    }
  }
  hipStream_t s = c->stream;
  rc = c->hip(hipMemcpyAsync(S->x, hx.data(), 24 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->vol, hv.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->m, hm.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->alpha, ha.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->beta, hb.data(), 8 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  S->hfixed.assign(3 * (size_t)N, 0);
  if (!rc) rc = c->hip(hipMemcpyAsync(S->fixed, S->hfixed.data(), 3 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (!rc) rc = thermal_sync(c);
  return rc;
}

// reference-order 3N host vector -> internal-order device vector
static int to_dev(Ctx* c, const double* h, double* d) {
  std::vector<double> t(3 * (size_t)c->N);
  for (int k = 0; k < c->N; k++)
    for (int q = 0; q < 3; q++) t[3 * (size_t)k + q] = h[3 * (size_t)c->perm[k] + q];
  int rc = c->hip(hipMemcpyAsync(d, t.data(), t.size() * 8, hipMemcpyHostToDevice, c->stream), "h2d");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

static int to_host(Ctx* c, const double* d, double* h) { return c->download_vec3(d, h); }

// device u -> residual in S->out (internal order)
static int residual(Ctx* c) {
  QsArgs a = args(c);
  hipStream_t s = c->stream;
  qs_theta<false, false>(a, s);
  qs_force<0>(a, s);
  return c->hip(hipGetLastError(), "qs residual");
}

// K w for device w (S->w), output in `out`; needs theta at the current u
static int tangent(Ctx* c, const double* w, double* out, bool theta_ready) {
  QsArgs a = args(c);
  a.w = w;
  a.out = out;
  hipStream_t s = c->stream;
  (void)theta_ready;  // the dual pass recomputes theta with its derivative
  qs_theta<true, false>(a, s);
  qs_force<1>(a, s);
  if (c->qs->damped) {
    const int n3 = 3 * c->N;
    hipLaunchKernelGGL(k_damp_add, dim3((n3 + RB - 1) / RB), dim3(RB), 0, s, n3, c->qs->dampv, w, out);
  }
  return c->hip(hipGetLastError(), "qs tangent");
}

static int block_jacobi(Ctx* c) {
  QsArgs a = args(c);
  hipStream_t s = c->stream;
  qs_theta<false, true>(a, s);
  qs_diag(a, s);
  return c->hip(hipGetLastError(), "qs block jacobi");
}

// Two-level preconditioner set-up at the current state: the coarse grid
// (at most coarse_points_cap(n) points, spacing equal on all axes, centred on
// the bounding box), the node lists of the restriction, and the inverse of
// the coarse matrix A_c = P^T K P, column by column from one tangent apply
// per coarse function. A_c is inverted on the host by Cholesky
// factorisation; a function whose pivot falls below 1e-12 of the largest
// diagonal entry (a grid point whose nodes' DOFs are all prescribed, or none
// at all) is dropped: its row and column of the inverse are 0.
static int coarse_build(Ctx* c) {
  QsState* S = c->qs;
  const int N = c->N;
  hipStream_t s = c->stream;
  auto A = [&](void** p, size_t bytes) { return c->hip(hipMalloc(p, std::max<size_t>(bytes, 64)), "qs coarse alloc"); };
  int rc = 0;
  const size_t mc = 3 * (size_t)kCoarsePointsMax;
  if (!S->clist) {
    rc = A((void**)&S->clist, 32 * (size_t)N);
    if (!rc) rc = A((void**)&S->wlist, 64 * (size_t)N);
    if (!rc) rc = A((void**)&S->cpart, 24 * (size_t)kCoarseSplit * kCoarsePointsMax);
    if (!rc) rc = A((void**)&S->cstart, 4 * (size_t)(kCoarsePointsMax + 1));
    if (!rc) rc = A((void**)&S->cinv, 8 * mc * mc);
    if (!rc) rc = A((void**)&S->cmat, 8 * mc * mc);
    if (!rc) rc = A((void**)&S->cy, 8 * mc);
    if (rc) return rc;
  }
  // grid over the bounding box
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int k = 0; k < N; k++)
    for (int d = 0; d < 3; d++) {
      const double v = c->x[3 * (size_t)k + d];
      lo[d] = std::min(lo[d], v);
      hi[d] = std::max(hi[d], v);
    }
  double E[3], emax = 0.0;
  for (int d = 0; d < 3; d++) { E[d] = std::max(hi[d] - lo[d], 0.0); emax = std::max(emax, E[d]); }
  if (!(emax > 0.0)) emax = 1.0;
  CoarseGrid G;
  for (int div = 1; div <= 64; div++) {
    const double H = emax / div;
    CoarseGrid T;
    T.H = H;
    for (int d = 0; d < 3; d++) T.d[d] = std::max(2, (int)std::ceil(E[d] / H - 1e-9) + 1);
    if (T.points() > coarse_points_cap(N)) break;
    for (int d = 0; d < 3; d++) T.o[d] = lo[d] - 0.5 * ((T.d[d] - 1) * H - E[d]);
    G = T;
  }
  const int np = G.points(), nc = 3 * np;
  S->cgr = G;
  S->nc = nc;
  // node lists of each grid point (the nodes of the up to 8 cells around it)
  std::vector<int> cell(N);
  const unsigned nb = (unsigned)((N + RB - 1) / RB);
  int* dcell = S->clist;  // scratch until the lists are written
  hipLaunchKernelGGL(k_coarse_cells, dim3(nb), dim3(RB), 0, s, N, G, S->x, dcell);
  rc = c->hip(hipMemcpyAsync(cell.data(), dcell, 4 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  std::vector<int> cnt(np + 1, 0), list(8 * (size_t)N);
  auto corner = [&](int cl, int q) {
    const int cx = cl % G.d[0], cy = (cl / G.d[0]) % G.d[1], cz = cl / (G.d[0] * G.d[1]);
    return (cx + (q & 1)) + G.d[0] * ((cy + ((q >> 1) & 1)) + G.d[1] * (cz + ((q >> 2) & 1)));
  };
  for (int i = 0; i < N; i++)
    for (int q = 0; q < 8; q++) cnt[corner(cell[i], q) + 1]++;
  for (int g = 0; g < np; g++) cnt[g + 1] += cnt[g];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < N; i++)
    for (int q = 0; q < 8; q++) list[fill[corner(cell[i], q)]++] = i;
  rc = c->hip(hipMemcpyAsync(S->cstart, cnt.data(), 4 * (size_t)(np + 1), hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipMemcpyAsync(S->clist, list.data(), 32 * (size_t)N, hipMemcpyHostToDevice, s), "h2d");
  if (rc) return rc;
  hipLaunchKernelGGL(k_coarse_weights, dim3(np), dim3(RB), 0, s, G, S->cstart, S->clist, S->x,
                     S->wlist);
  // A_c by probing: the tangent couples nodes at most 2 bond lengths apart
  // (the dilatation term reaches the neighbours' neighbours), so phi_g' . K
  // phi_g vanishes unless g and g' are within `reach` grid steps on every
  // axis (supports H either side of a point, bond components at most
  // bond_cmax). Points of one colour (indices congruent modulo 2 reach + 1 on
  // every axis) then never meet the same target, and one tangent apply of
  // their summed functions gives all their columns at once. S->z holds the
  // probe, S->q its image; probe p's restriction goes to column p of S->cmat.
  const double bc = c->bond_cmax;
  const int reach = std::max(1, (int)std::ceil((2.0 * G.H + 2.0 * bc) / G.H - 1e-12) - 1);
  const int stride = 2 * reach + 1;
  const int sc[3] = {std::min(stride, G.d[0]), std::min(stride, G.d[1]), std::min(stride, G.d[2])};
  std::vector<int> pcol;  // probe -> colour code (cx, cy, cz, k0)
  for (int cz = 0; cz < sc[2]; cz++)
    for (int cy = 0; cy < sc[1]; cy++)
      for (int cx = 0; cx < sc[0]; cx++)
        for (int k0 = 0; k0 < 3; k0++) {
          const int pr = (int)pcol.size();
          pcol.push_back(((cz * sc[1] + cy) * sc[0] + cx) * 3 + k0);
          hipLaunchKernelGGL(k_coarse_probe, dim3(nb), dim3(RB), 0, s, N, G, stride, cx, cy, cz, k0,
                             S->x, S->fixed, S->z);
          if ((rc = tangent(c, S->z, S->q, true))) return rc;
          hipLaunchKernelGGL(k_coarse_partial, dim3(np * kCoarseSplit), dim3(RB), 0, s, S->cstart,
                             S->clist, S->wlist, S->fixed, S->q, S->cpart);
          hipLaunchKernelGGL(k_coarse_finish<false>, dim3(1), dim3(RB), 0, s, nc, S->cpart, nullptr,
                             S->cmat + (size_t)pr * nc);
        }
  const int nprobe = (int)pcol.size();  // <= nc, so S->cmat holds every probe
  std::vector<double> R((size_t)nprobe * nc), M((size_t)nc * nc, 0.0);
  rc = c->hip(hipGetLastError(), "qs coarse build");
  if (!rc) rc = c->hip(hipMemcpyAsync(R.data(), S->cmat, 8 * R.size(), hipMemcpyDeviceToHost, s), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  // entry (target t, source g): the source of the target's colour within
  // reach on every axis, if any (M[col * nc + row] = A_c[row][col])
  for (int pr = 0; pr < nprobe; pr++) {
    const int k0 = pcol[pr] % 3, code = pcol[pr] / 3;
    const int cx = code % sc[0], cy = (code / sc[0]) % sc[1], cz = code / (sc[0] * sc[1]);
    const int cc[3] = {cx, cy, cz};
    for (int t = 0; t < np; t++) {
      const int tp[3] = {t % G.d[0], (t / G.d[0]) % G.d[1], t / (G.d[0] * G.d[1])};
      int gp[3];
      bool ok = true;
      for (int d = 0; d < 3 && ok; d++) {
        int delta = ((cc[d] - tp[d]) % stride + stride) % stride;
        if (delta > reach) delta -= stride;
        gp[d] = tp[d] + delta;
        ok = gp[d] >= 0 && gp[d] < G.d[d] && gp[d] % stride == cc[d];
      }
      if (!ok) continue;
      const int g = gp[0] + G.d[0] * (gp[1] + G.d[1] * gp[2]);
      for (int k = 0; k < 3; k++) M[(size_t)(3 * g + k0) * nc + 3 * t + k] = R[(size_t)pr * nc + 3 * t + k];
    }
  }
  // symmetrise, factor L L^T with dropped functions, invert
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < i; j++) {
      const double v = 0.5 * (M[(size_t)i * nc + j] + M[(size_t)j * nc + i]);
      M[(size_t)i * nc + j] = M[(size_t)j * nc + i] = v;
    }
  double dmax = 0.0;
  for (int i = 0; i < nc; i++) dmax = std::max(dmax, M[(size_t)i * nc + i]);
  std::vector<double> L((size_t)nc * nc, 0.0);
  std::vector<char> drop(nc, 0);
  for (int j = 0; j < nc; j++) {
    double d = M[(size_t)j * nc + j];
    for (int k = 0; k < j; k++) d -= L[(size_t)j * nc + k] * L[(size_t)j * nc + k];
    if (!(d > 1e-12 * dmax)) { drop[j] = 1; continue; }
    const double ljj = std::sqrt(d);
    L[(size_t)j * nc + j] = ljj;
    for (int i = j + 1; i < nc; i++) {
      double v = M[(size_t)i * nc + j];
      for (int k = 0; k < j; k++) v -= L[(size_t)i * nc + k] * L[(size_t)j * nc + k];
      L[(size_t)i * nc + j] = v / ljj;
    }
  }
  std::vector<double> inv((size_t)nc * nc, 0.0);
  parallel_for(nc, [&](long long e0, long long e1) {  // columns on host threads
  std::vector<double> y(nc);
  for (int e = (int)e0; e < (int)e1; e++) {
    if (drop[e]) continue;
    for (int i = 0; i < nc; i++) {  // L y = e_e
      if (drop[i]) { y[i] = 0.0; continue; }
      double v = i == e ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) v -= L[(size_t)i * nc + k] * y[k];
      y[i] = v / L[(size_t)i * nc + i];
    }
    for (int i = nc - 1; i >= 0; i--) {  // L^T x = y
      if (drop[i]) { y[i] = 0.0; continue; }
      double v = y[i];
      for (int k = i + 1; k < nc; k++) v -= L[(size_t)k * nc + i] * y[k];
      y[i] = v / L[(size_t)i * nc + i];
    }
    for (int i = 0; i < nc; i++) inv[(size_t)i * nc + e] = y[i];
  }
  }, 64);
  rc = c->hip(hipMemcpyAsync(S->cinv, inv.data(), 8 * inv.size(), hipMemcpyHostToDevice, s), "h2d");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  S->coarse_ok = rc == 0;
  return rc;
}

static int dot(Ctx* c, const double* x, const double* y, double* out) {
  QsState* S = c->qs;
  const int n = 3 * c->N;
  hipLaunchKernelGGL(k_dot, dim3(S->nblk), dim3(RB), 0, c->stream, n, x, y, S->part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, c->stream, S->nblk, S->part, 1, S->scal);
  int rc = c->hip(hipMemcpyAsync(out, S->scal, 8, hipMemcpyDeviceToHost, c->stream), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

// ||v||_2 and ||v||_inf over all entries (mask null) or the masked ones.
static int norms(Ctx* c, const double* v, const signed char* mask, double* n2, double* ninf) {
  QsState* S = c->qs;
  const int n = 3 * c->N;
  hipLaunchKernelGGL(k_norm_part, dim3(S->nblk), dim3(RB), 0, c->stream, n, v, mask, S->part);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(RB), 0, c->stream, S->nblk, S->part, S->scal + kScal - 2);
  double h[2] = {0.0, 0.0};
  int rc = c->hip(hipMemcpyAsync(h, S->scal + kScal - 2, 16, hipMemcpyDeviceToHost, c->stream), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  *n2 = std::sqrt(h[0]);
  *ninf = h[1];
  return rc;
}

// Residual at the device state S->u into S->out and its norm
// ||r||_2 + 20 ||r||_inf (src/core/Peridigm.cpp:4306-4316).
static int residual_norm(Ctx* c, double* norm) {
  int rc = residual(c);
  double n2 = 0.0, ninf = 0.0;
  if (!rc) rc = norms(c, c->qs->out, nullptr, &n2, &ninf);
  *norm = n2 + 20.0 * ninf;
  return rc;
}

// ||f V||_2 over the kinematic DOFs at S->u: the reaction norm that scales
// the Newton tolerance (src/core/Peridigm.cpp:2716-2732). Leaves S->out
// holding f V on every DOF.
static int reaction_norm(Ctx* c, double* norm) {
  QsArgs a = args(c);
  hipStream_t s = c->stream;
  qs_theta<false, false>(a, s);
  qs_force<2>(a, s);
  double ninf = 0.0;
  int rc = c->hip(hipGetLastError(), "qs reactions");
  if (!rc) rc = norms(c, c->qs->out, c->qs->fixed, norm, &ninf);
  return rc;
}

// Destroys the cached iteration graphs (after the stream has drained).
static void drop_graphs(Ctx* c) {
  QsState* S = c->qs;
  if (!S) return;
  bool any = false;
  for (auto& a : S->gexec) for (auto& b : a) for (hipGraphExec_t g : b) any = any || g;
  if (!any) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& a : S->gexec)
    for (auto& b : a)
      for (hipGraphExec_t& g : b)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

}  // namespace qs

int qs_tangent_apply(Ctx* c, const double* w, double* out) { return qs::tangent(c, w, out, true); }
int qs_dot(Ctx* c, const double* x, const double* y, double* out) { return qs::dot(c, x, y, out); }

double* Ctx::qs_dinv() { return qs ? qs->dinv : nullptr; }
const signed char* Ctx::qs_fixed() { return qs ? qs->fixed : nullptr; }

int pdg_qs_prepare(Ctx* c, const double* u) {
  int rc = qs::ensure(c);
  if (!rc) rc = qs::to_dev(c, u, c->qs->u);
  if (!rc) rc = qs::block_jacobi(c);
  return rc;
}

void Ctx::free_extra() {
  for (hipEvent_t e : tev) (void)hipEventDestroy(e);
  tev.clear();
  if (qs) {
    QsState* S = qs;
    double* ptrs[] = {S->x,   S->u,  S->w,    S->vol,  S->m,      S->alpha, S->beta,
                      S->theta, S->dtheta, S->g, S->out, S->dinv, S->xs,    S->r,
                      S->z,   S->p,  S->q,    S->b,    S->part,   S->scal,  S->diag,
                      S->dampv, S->dT,   S->ps};
    for (double* p : ptrs) if (p) (void)hipFree(p);
    for (double* p : S->rv) if (p) (void)hipFree(p);
    double* cptrs[] = {S->cinv, S->cmat, S->cy, S->wlist, S->cpart};
    for (double* p : cptrs) if (p) (void)hipFree(p);
    if (S->clist) (void)hipFree(S->clist);
    if (S->cstart) (void)hipFree(S->cstart);
    for (int k = 0; k < 4; k++) {
      if (S->ph[k]) (void)hipFree(S->ph[k]);
      if (S->qh[k]) (void)hipFree(S->qh[k]);
    }
    qs::drop_graphs(this);
    if (S->fixed) (void)hipFree(S->fixed);
    if (S->hscal) (void)hipHostFree(S->hscal);
    delete S;
    qs = nullptr;
  }
  free_gnn();
  free_contact();  // sized by the topology and horizon
}

// K x = b on device vectors in internal order; x holds the initial guess.
// Returns iterations in *iters, the final relative residual in *rel.
// Preconditioned CG (precond 0: none, 1: block Jacobi, 3: block Jacobi plus
// the coarse correction, built by coarse_build) or, for the learned
// preconditioner (precond 2), Notay's flexible CG truncated to the last
// kFcgSlots - 1 directions: the new direction is z minus its K-projections on
// them, alpha = p.r / p.Kp. Scalars stay on the device; the host reads the
// residual norm at checks only: every iteration for the first 8, then every 4
// (a solve may run up to 3 iterations past the tolerance).
int qs_solve(Ctx* c, double* x, const double* b, int precond, double tol, int maxit, int* iters,
             double* rel) {
  QsState* S = c->qs;
  const int n3 = 3 * c->N;
  hipStream_t s = c->stream;
  const unsigned vb = (unsigned)((n3 + RB - 1) / RB);
  const unsigned nb = (unsigned)((c->N + RB - 1) / RB);
  const dim3 red(S->nblk), blk(RB);
  const bool flexible = precond == 2;
  int rc = 0;
  double bb = 0.0;
  if ((rc = qs::dot(c, b, b, &bb))) return rc;
  const double bnorm = std::sqrt(bb);
  *iters = 0;
  if (bnorm == 0.0) {
    rc = c->hip(hipMemsetAsync(x, 0, 8 * (size_t)n3, s), "zero x");
    *rel = 0.0;
    return rc;
  }
  // device dot product into scalar slot k
  auto ddot = [&](const double* u, const double* v, int k) {
    hipLaunchKernelGGL(k_dot, red, blk, 0, s, n3, u, v, S->part);
    hipLaunchKernelGGL(k_final, dim3(1), blk, 0, s, S->nblk, S->part, 1, S->scal + k);
  };
  // coarse part of the two-level preconditioner: y = A_c^-1 P^T r
  auto coarse = [&](const double* r) -> int {
    hipLaunchKernelGGL(k_coarse_partial, dim3(S->cgr.points() * kCoarseSplit), blk, 0, s,
                       S->cstart, S->clist, S->wlist, S->fixed, r, S->cpart);
    hipLaunchKernelGGL(k_coarse_finish<true>, dim3((S->nc + RB / 64 - 1) / (RB / 64)), blk, 0, s,
                       S->nc, S->cpart, S->cinv, S->cy);
    return c->hip(hipGetLastError(), "coarse solve");
  };
  auto precond_apply = [&](const double* r, double* z) -> int {
    if (precond == 2) return gnn_apply(c, r, z);
    if (precond == 3) {
      int r2 = coarse(r);
      if (!r2) hipLaunchKernelGGL((k_two_level_apply<false>), dim3(nb), blk, 0, s, c->N, S->dinv,
                                  S->cgr, S->x, S->fixed, S->cy, r, z, nullptr);
      return r2 ? r2 : c->hip(hipGetLastError(), "two-level apply");
    }
    if (precond == 1) {
      hipLaunchKernelGGL(k_block_apply, dim3(nb), blk, 0, s, c->N, S->dinv, r, z);
      return c->hip(hipGetLastError(), "block apply");
    }
    return c->hip(hipMemcpyAsync(z, r, 8 * (size_t)n3, hipMemcpyDeviceToDevice, s), "copy");
  };
  // r = b - K x; z = M r; p = z
  int cur = 0;  // flexible CG: history slot of the current direction
  double* p = flexible ? S->ph[0] : S->p;
  if ((rc = qs::tangent(c, x, S->q, false))) return rc;
  hipLaunchKernelGGL(k_sub, dim3(vb), blk, 0, s, n3, b, S->q, S->r);
  rc = c->hip(hipMemsetAsync(S->scal, 0, 8 * kScal, s), "zero scalars");
  if (!rc) rc = precond_apply(S->r, S->z);
  if (!rc) rc = c->hip(hipMemcpyAsync(p, S->z, 8 * (size_t)n3, hipMemcpyDeviceToDevice, s), "copy");
  if (rc) return rc;
  ddot(S->r, S->z, SC_RZN);  // r.z of the first iteration (read by k_cg_step)
  ddot(S->r, S->r, SC_RR);
  auto check = [&](double* relres, bool* brk) -> int {
    int r2 = c->hip(hipMemcpyAsync(S->hscal, S->scal, 8 * kScal, hipMemcpyDeviceToHost, s), "d2h");
    if (!r2) r2 = c->hip(hipStreamSynchronize(s), "sync");
    *relres = std::sqrt(S->hscal[SC_RR]) / bnorm;
    *brk = S->hscal[SC_BRK] != 0.0;
    return r2;
  };
  double relres = 0.0;
  bool brk = false;
  if ((rc = check(&relres, &brk))) return rc;
  int it = 0, nhist = 0;
  // stagnation watch for the learned preconditioner: every 48 iterations the
  // residual must have halved, else flexible CG hands over to block-Jacobi CG
  double watch = relres;
  bool stalled = false;
  // The iteration is two fixed launch sequences: A (K p, p.Kp, step, r.r)
  // and B (preconditioner, dots, new direction), split by the host check.
  // For flexible CG both depend on the history slot `cur`; once the history
  // is full (nhist = kFcgSlots - 1) they are the same every time that slot
  // comes round. With graphs on, each (preconditioner, sequence, slot) is
  // captured once (kept in the solver state until its vectors or the
  // network's buffers change) and replayed.
  // rr: also reduce r.r to its slot (needed only where the host checks it)
  auto seq_a = [&](int cs, bool rr) -> int {
    double* pp = flexible ? S->ph[cs] : S->p;
    double* qq = flexible ? S->qh[cs] : S->q;
    int r2 = qs::tangent(c, pp, qq, true);
    if (r2) return r2;
    double* rrp = S->part;  // r.r partials
    if (flexible) {
      ddot(pp, qq, SC_PQH + cs);
      ddot(pp, S->r, SC_PR);
      hipLaunchKernelGGL(k_fcg_step, red, blk, 0, s, n3, S->scal, cs, pp, qq, x, S->r, S->part);
    } else {
      hipLaunchKernelGGL(k_dot, red, blk, 0, s, n3, pp, qq, S->part);
      rrp = S->part + S->nblk;
      hipLaunchKernelGGL(k_cg_step, red, blk, 0, s, n3, S->nblk, S->scal, S->part, pp, qq, x, S->r,
                         rrp);
    }
    if (rr) hipLaunchKernelGGL(k_final, dim3(1), blk, 0, s, S->nblk, rrp, 1, S->scal + SC_RR);
    return c->hip(hipGetLastError(), "cg step");
  };
  auto seq_b = [&](int cs, int nh) -> int {
    int r2 = 0;
    if (flexible) {
      r2 = precond_apply(S->r, S->z);
    } else if (precond == 1) {
      // z = Dinv r with the r.z partials in the same pass
      hipLaunchKernelGGL(k_block_apply_dot, red, blk, 0, s, c->N, S->dinv, S->r, S->z, S->part);
    } else if (precond == 3) {
      r2 = coarse(S->r);
      if (!r2) hipLaunchKernelGGL((k_two_level_apply<true>), red, blk, 0, s, c->N, S->dinv, S->cgr,
                                  S->x, S->fixed, S->cy, S->r, S->z, S->part);
    } else {
      r2 = precond_apply(S->r, S->z);
      if (!r2) hipLaunchKernelGGL(k_dot, red, blk, 0, s, n3, S->r, S->z, S->part);
    }
    if (r2) return r2;
    if (flexible) {
      FcgHist h{};
      h.nh = nh;
      for (int k = 0; k < nh; k++) {
        const int slot = (cs - k + kFcgSlots) % kFcgSlots;
        h.p[k] = S->ph[slot];
        h.slot[k] = slot;
        ddot(S->z, S->qh[slot], SC_ZQH + slot);
      }
      hipLaunchKernelGGL(k_fcg_dir, dim3(vb), blk, 0, s, n3, S->scal, S->z, h,
                         S->ph[(cs + 1) % kFcgSlots]);
    } else {
      hipLaunchKernelGGL(k_cg_dir, red, blk, 0, s, n3, S->nblk, S->scal, S->part, S->z, S->p);
    }
    return c->hip(hipGetLastError(), "cg direction");
  };
  // the damped tangent adds a launch: no graphs while damping
  const bool graphs = S->graphs && !S->damped && precond < 3;
  // graphs captured for other vectors or older network buffers are stale
  if (graphs && (S->gx != x || S->gb != b || (flexible && S->gepoch != c->gnn_epoch))) {
    qs::drop_graphs(c);
    S->gx = x; S->gb = b; S->gepoch = c->gnn_epoch;
  }
  hipGraphExec_t* ga = S->gexec[std::min(precond, 2)][0];
  hipGraphExec_t* gb = S->gexec[std::min(precond, 2)][1];
  auto capture = [&](hipGraphExec_t* out, auto&& body) -> int {
    hipGraph_t g = nullptr;
    int r2 = c->hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    if (r2) return r2;
    const int rb = body();
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rb) { if (g) (void)hipGraphDestroy(g); return rb; }
    r2 = c->hip(e, "end capture");
    if (!r2) r2 = c->hip(hipGraphInstantiate(out, g, nullptr, nullptr, 0), "instantiate");
    if (g) (void)hipGraphDestroy(g);
    return r2;
  };
  // runs sequence A or B for slot cs: replayed from a graph when the history
  // is full and graphs are on, otherwise launched directly
  auto run_a = [&](int cs, bool full, bool rr) -> int {
    if (!graphs || !full) return seq_a(cs, rr);
    // a captured sequence is replayed before every check: it reduces r.r
    if (!ga[cs]) { int r2 = capture(&ga[cs], [&]() { return seq_a(cs, true); }); if (r2) return r2; }
    return c->hip(hipGraphLaunch(ga[cs], s), "graph launch");
  };
  auto run_b = [&](int cs, int nh, bool full) -> int {
    if (!graphs || !full) return seq_b(cs, nh);
    if (!gb[cs]) { int r2 = capture(&gb[cs], [&]() { return seq_b(cs, nh); }); if (r2) return r2; }
    return c->hip(hipGraphLaunch(gb[cs], s), "graph launch");
  };
  const int hist_full = flexible ? kFcgSlots - 1 : 0;
  // the host checks the residual after iterations 1-8, every fourth, and the last
  auto checked = [&](int k) { return k <= 8 || k % 4 == 0 || k == maxit; };
  while (relres > tol && it < maxit) {
    if ((rc = run_a(cur, nhist >= hist_full, checked(it + 1)))) return rc;
    it++;
    if (checked(it)) {
      if ((rc = check(&relres, &brk))) return rc;
      if (brk) return c->fail(PDG_E_NOCONV, "CG breakdown: p.Kp <= 0");
      if (relres <= tol) break;
      if (flexible && it % 48 == 0) {
        if (!(relres < 0.5 * watch)) { stalled = true; break; }
        watch = relres;
      }
    }
    if (flexible) {
      const int nh = std::min(nhist + 1, kFcgSlots - 1);
      if ((rc = run_b(cur, nh, nh >= hist_full && nhist >= hist_full))) return rc;
      nhist = nh;
      cur = (cur + 1) % kFcgSlots;
    } else {
      if ((rc = run_b(0, 0, true))) return rc;
    }
  }
  if (!stalled && relres > tol) {
    if ((rc = check(&relres, &brk))) return rc;
    if (brk) return c->fail(PDG_E_NOCONV, "CG breakdown: p.Kp <= 0");
  }
  if (stalled) {
    // fallback: block-Jacobi CG from the current iterate
    c->qs->fallbacks++;
    int it2 = 0;
    rc = qs_solve(c, x, b, 1, tol, std::max(0, maxit - it), &it2, rel);
    *iters = it + it2;
    return rc;
  }
  *iters = it;
  *rel = relres;
  if (relres > tol) return c->fail(PDG_E_NOCONV, "linear solve did not converge");
  return c->hip(hipGetLastError(), "cg");
}

}  // namespace pdg

using namespace pdg;

extern "C" {

int pdg_qs_set_fixed(pdg_ctx* h, const int* dofs, int numDofs) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || numDofs < 0 || (numDofs > 0 && !dofs)) return PDG_E_ARG;
  int rc = qs::ensure(c);
  if (rc) return rc;
  QsState* S = c->qs;
  std::fill(S->hfixed.begin(), S->hfixed.end(), 0);
  for (int k = 0; k < numDofs; k++) {
    const int d = dofs[k];
    if (d < 0 || d >= 3 * c->N) return c->fail(PDG_E_ARG, "fixed DOF out of range");
    S->hfixed[3 * (size_t)c->iperm[d / 3] + d % 3] = 1;
  }
  rc = c->hip(hipMemcpyAsync(S->fixed, S->hfixed.data(), S->hfixed.size(), hipMemcpyHostToDevice, c->stream), "h2d");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  S->coarse_ok = false;  // the coarse functions vanish on prescribed DOFs
  return rc;
}

int pdg_qs_residual(pdg_ctx* h, const double* u, double* residual, double* norm) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u) return PDG_E_ARG;
  int rc = qs::ensure(c);
  if (!rc) rc = qs::to_dev(c, u, c->qs->u);
  if (!rc) rc = qs::residual(c);
  if (rc) return rc;
  std::vector<double> r(3 * (size_t)c->N);
  rc = qs::to_host(c, c->qs->out, r.data());
  if (rc) return rc;
  double n2 = 0.0, ninf = 0.0;
  for (double v : r) { n2 += v * v; ninf = std::max(ninf, std::fabs(v)); }
  if (residual) std::copy(r.begin(), r.end(), residual);
  if (norm) *norm = std::sqrt(n2) + 20.0 * ninf;
  return PDG_OK;
}

int pdg_qs_set_graphs(pdg_ctx* h, int enable) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  int rc = qs::ensure(c);
  if (!rc) c->qs->graphs = enable != 0;
  return rc;
}

int pdg_qs_fallback_count(pdg_ctx* h, int* count) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !count) return PDG_E_ARG;
  *count = c->qs ? c->qs->fallbacks : 0;
  return PDG_OK;
}

int pdg_qs_tangent_apply(pdg_ctx* h, const double* u, const double* w, double* Kw) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || !w || !Kw) return PDG_E_ARG;
  int rc = qs::ensure(c);
  if (!rc) rc = qs::to_dev(c, u, c->qs->u);
  if (!rc) rc = qs::to_dev(c, w, c->qs->w);
  if (!rc) rc = qs::tangent(c, c->qs->w, c->qs->out, false);
  if (!rc) rc = qs::to_host(c, c->qs->out, Kw);
  return rc;
}

int pdg_qs_block_diagonal(pdg_ctx* h, const double* u, double* blocksInv) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || !blocksInv) return PDG_E_ARG;
  int rc = qs::ensure(c);
  if (!rc) rc = qs::to_dev(c, u, c->qs->u);
  if (!rc) rc = qs::block_jacobi(c);
  if (rc) return rc;
  std::vector<double> t(9 * (size_t)c->N);
  rc = c->hip(hipMemcpyAsync(t.data(), c->qs->dinv, t.size() * 8, hipMemcpyDeviceToHost, c->stream), "d2h");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  for (int k = 0; k < c->N; k++)
    for (int q = 0; q < 9; q++) blocksInv[9 * (size_t)c->perm[k] + q] = t[9 * (size_t)k + q];
  return PDG_OK;
}

int pdg_qs_solve(pdg_ctx* h, const double* u, const double* rhs, double* x, int precond,
                 double relTol, int maxIter, int* iters, double* relRes) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || !rhs || !x || precond < 0 || precond > 3) return PDG_E_ARG;
  int rc = qs::ensure(c);
  QsState* S = c->qs;
  if (!rc) rc = qs::to_dev(c, u, S->u);
  if (!rc && precond >= 1) rc = qs::block_jacobi(c);
  if (!rc && precond == 3 && !S->coarse_ok) rc = qs::coarse_build(c);
  if (!rc) rc = qs::to_dev(c, rhs, S->b);
  if (!rc) rc = qs::to_dev(c, x, S->xs);
  if (rc) return rc;
  int it = 0;
  double rel = 0.0;
  rc = qs_solve(c, S->xs, S->b, precond, relTol, maxIter, &it, &rel);
  if (iters) *iters = it;
  if (relRes) *relRes = rel;
  int rc2 = qs::to_host(c, S->xs, x);
  return rc ? rc : rc2;
}

// Newton iterations at fixed prescribed values: u (reference order) holds
// the prescribed values on fixed DOFs and the initial guess elsewhere.
// Converged when ||r||_2 + 20 ||r||_inf <= tol * ||reactions||_2, with the
// reactions (f V on the kinematic DOFs) computed once, at the entry state,
// as the reference does once per load step (src/core/Peridigm.cpp:2700-2732);
// with the absolute-tolerance flag set the multiplier is 1. Each linear
// solve runs to linTol.
int pdg_qs_newton(pdg_ctx* h, double* u, double tol, int maxNewton, int precond, double linTol,
                  int maxLin, int* newtonIters, int* linIters, double* residualNorm) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || precond < 0 || precond > 3) return PDG_E_ARG;
  int rc = qs::ensure(c);
  QsState* S = c->qs;
  if (!rc) rc = qs::to_dev(c, u, S->u);
  if (rc) return rc;
  const int n3 = 3 * c->N;
  const unsigned vb = (unsigned)((n3 + RB - 1) / RB);
  int nit = 0, lit = 0;
  bool converged = false;
  double rnorm = 0.0;
  std::vector<double> hr(n3);
  double multiplier = 1.0;
  if (!c->qs_abs_tol) {
    // reactions: the residual with no rows zeroed, read on the kinematic DOFs
    std::vector<double> hreac(n3);
    std::vector<signed char> none(n3, 0);
    rc = c->hip(hipMemcpyAsync(S->fixed, none.data(), n3, hipMemcpyHostToDevice, c->stream), "h2d");
    if (!rc) rc = qs::residual(c);
    if (!rc) rc = c->hip(hipMemcpyAsync(hreac.data(), S->out, 8 * (size_t)n3, hipMemcpyDeviceToHost, c->stream), "d2h");
    if (!rc) rc = c->hip(hipMemcpyAsync(S->fixed, S->hfixed.data(), n3, hipMemcpyHostToDevice, c->stream), "h2d");
    if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    double reac = 0.0;
    for (int q = 0; q < n3; q++) if (S->hfixed[q]) reac += hreac[q] * hreac[q];
    multiplier = std::sqrt(reac);
  }
  c->qs_tol_multiplier = multiplier;
  for (;; nit++) {
    if ((rc = qs::residual(c))) return rc;
    rc = c->hip(hipMemcpyAsync(hr.data(), S->out, 8 * (size_t)n3, hipMemcpyDeviceToHost, c->stream), "d2h");
    if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    double n2 = 0.0, ninf = 0.0;
    for (double v : hr) { n2 += v * v; ninf = std::max(ninf, std::fabs(v)); }
    rnorm = std::sqrt(n2) + 20.0 * ninf;
    if (rnorm <= tol * multiplier) { converged = true; break; }
    if (nit == maxNewton) break;
    // the learned and two-level preconditioners use the diagonal blocks too
    if (precond >= 1 && (rc = qs::block_jacobi(c))) return rc;
    if (precond == 3 && !S->coarse_ok && (rc = qs::coarse_build(c))) return rc;
    rc = c->hip(hipMemcpyAsync(S->b, S->out, 8 * (size_t)n3, hipMemcpyDeviceToDevice, c->stream), "copy");
    if (!rc) rc = c->hip(hipMemsetAsync(S->xs, 0, 8 * (size_t)n3, c->stream), "zero");
    if (rc) return rc;
    int it = 0;
    double rel = 0.0;
    rc = qs_solve(c, S->xs, S->b, precond, linTol, maxLin, &it, &rel);
    lit += it;
    if (rc) return rc;
    hipLaunchKernelGGL(k_axpby, dim3(vb), dim3(RB), 0, c->stream, n3, 1.0, S->xs, 1.0, S->u);
  }
  if (newtonIters) *newtonIters = nit;
  if (linIters) *linIters = lit;
  if (residualNorm) *residualNorm = rnorm;
  rc = qs::to_host(c, S->u, u);
  if (rc) return rc;
  if (!converged) return c->fail(PDG_E_NOCONV, "Newton did not converge");
  return PDG_OK;
}

int pdg_qs_force_dilatation(pdg_ctx* h, const double* u, double* force, double* dilatation) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || !force) return PDG_E_ARG;
  (void)hipSetDevice(c->device);
  int rc = qs::ensure(c);
  QsState* S = c->qs;
  if (!rc) rc = qs::to_dev(c, u, S->u);
  if (rc) return rc;
  QsArgs a = qs::args(c);
  hipStream_t s = c->stream;
  qs_theta<false, false>(a, s);
  qs_force<2>(a, s);
  rc = c->hip(hipGetLastError(), "qs force");
  if (!rc) rc = qs::to_host(c, S->out, force);
  if (rc) return rc;
  for (int p = 0; p < c->N; p++)
    for (int d = 0; d < 3; d++) force[3 * (size_t)p + d] /= c->vol[p];
  if (dilatation) {
    std::vector<double> th(c->N);
    rc = c->hip(hipMemcpyAsync(th.data(), S->theta, 8 * (size_t)c->N, hipMemcpyDeviceToHost, s), "d2h");
    if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    for (int k = 0; k < c->N; k++) dilatation[c->perm[k]] = th[k];
  }
  return PDG_OK;
}

int pdg_set_thermal_expansion(pdg_ctx* h, double coefficient) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !std::isfinite(coefficient)) return PDG_E_ARG;
  if (coefficient != 0.0 && c->material != PDG_MATERIAL_ELASTIC)
    return c->fail(PDG_E_UNSUPPORTED, "thermal strains belong to the elastic (LPS) model");
  c->alpha_t = coefficient;
  c->thermal_epoch++;
  if (c->qs) qs::drop_graphs(c);
  return PDG_OK;
}

int pdg_set_temperature_change(pdg_ctx* h, const double* dT0, const double* dT1) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts) return PDG_E_ARG;
  const int N = c->N;
  if (!dT0 && !dT1) {
    c->dT0.clear();
    c->dT1.clear();
  } else {
    c->dT0.assign(N, 0.0);
    if (dT0) c->dT0.assign(dT0, dT0 + N);
    if (dT1) c->dT1.assign(dT1, dT1 + N);
    else c->dT1.clear();
  }
  c->thermal_epoch++;
  if (c->qs) qs::drop_graphs(c);
  return PDG_OK;
}

int pdg_set_temperature_time(pdg_ctx* h, double t) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !std::isfinite(t)) return PDG_E_ARG;
  c->temp_time = t;
  return PDG_OK;
}

int pdg_compute_partial_stress(pdg_ctx* h, const double* y, const double* bondDamage,
                               double* partialStress) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || (c->N > 0 && (!y || !partialStress))) return PDG_E_ARG;
  if (c->material != PDG_MATERIAL_ELASTIC)
    return c->fail(PDG_E_UNSUPPORTED, "the partial stress belongs to the elastic (LPS) model");
  (void)hipSetDevice(c->device);
  int rc = qs::ensure(c);
  if (rc) return rc;
  QsState* S = c->qs;
  const int N = c->N;
  if (!S->ps && (rc = c->hip(hipMalloc((void**)&S->ps, std::max<size_t>(72 * (size_t)N, 64)), "qs alloc")))
    return rc;
  if ((rc = c->upload_mask(bondDamage, true))) return rc;
  std::vector<double> u(3 * (size_t)N);
  for (size_t q = 0; q < u.size(); q++) u[q] = y[q] - c->x[q];
  if ((rc = qs::to_dev(c, u.data(), S->u))) return rc;
  QsArgs a = qs::args(c);
  hipStream_t s = c->stream;
  qs_theta<false, false>(a, s);
  qs_force<3>(a, s);
  rc = c->hip(hipGetLastError(), "partial stress");
  std::vector<double> hp(9 * (size_t)N);
  if (!rc) rc = c->hip(hipMemcpyAsync(hp.data(), S->ps, 72 * (size_t)N, hipMemcpyDeviceToHost, s), "d2h ps");
  if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
  if (rc) return rc;
  for (int k = 0; k < N; k++)
    for (int t = 0; t < 9; t++) partialStress[9 * (size_t)c->perm[k] + t] = hp[9 * (size_t)k + t];
  return PDG_OK;
}

int pdg_qs_set_absolute_tolerance(pdg_ctx* h, int on) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return PDG_E_ARG;
  c->qs_abs_tol = on != 0;
  return PDG_OK;
}

int pdg_qs_tolerance_multiplier(const pdg_ctx* h, double* multiplier) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c || !multiplier) return PDG_E_ARG;
  *multiplier = c->qs_tol_multiplier;
  return PDG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Quasi-static load stepping (Peridigm::executeQuasiStatic,
// src/core/Peridigm.cpp:2470-3043): per load step the prescribed-displacement
// increments, the reaction-scaled tolerance, the Newton loop with the
// previous step's velocity as predictor, the damped-Newton switch and retry,
// the candidate-alpha line search (:3364-3618), and adaptive halving of the
// load step (:2889-2934). The tangent is applied matrix-free at the current
// state in every iteration; the reference reassembles it every iteration
// too, except in damped Newton, where it keeps one for 8 iterations.
// ---------------------------------------------------------------------------
namespace pdg {
namespace {

__global__ __launch_bounds__(RB) void k_div(int n, const double* x, double d, double* y) {
  const int i = blockIdx.x * RB + threadIdx.x;
  if (i < n) y[i] = x[i] / d;
}

struct QsRun {
  Ctx* c;
  QsState* S;
  int n3;
  unsigned vb;
  double *ub, *du, *pred, *tmp, *rs;  // S->rv
  double* lhs;                        // S->xs

  // S->u = ub + du + a * lhs (the reference's y = x + u + deltaU)
  void state(double a) {
    hipStream_t s = c->stream;
    if (a == 0.0) {
      hipLaunchKernelGGL(k_lin2, dim3(vb), dim3(RB), 0, s, n3, 1.0, ub, 1.0, du, S->u);
    } else {
      hipLaunchKernelGGL(k_lin2, dim3(vb), dim3(RB), 0, s, n3, 1.0, du, a, lhs, tmp);
      hipLaunchKernelGGL(k_lin2, dim3(vb), dim3(RB), 0, s, n3, 1.0, ub, 1.0, tmp, S->u);
    }
  }

  // residual norm with the increment du + a * lhs
  int trial(double a, double* norm) {
    state(a);
    return qs::residual_norm(c, norm);
  }

  // quasiStaticsLineSearch (src/core/Peridigm.cpp:3364-3618)
  int line_search(double* alpha_out) {
    double r0 = 0.0;
    int rc = trial(0.0, &r0);
    if (rc) return rc;
    if (!std::isfinite(r0)) return c->fail(PDG_E_NONFINITE, "NaN in the residual of the line search");
    if (r0 == 0.0) { *alpha_out = 0.0; return PDG_OK; }
    rc = c->hip(hipMemcpyAsync(rs, S->out, 8 * (size_t)n3, hipMemcpyDeviceToDevice, c->stream), "copy");
    if (rc) return rc;
    double best = 1.0, bestres = 1.0e50;
    std::vector<double> cand;
    // the secant estimate from a small perturbation
    const double eps = 1.0e-4;
    double rp = 0.0, SR = 0.0, SPR = 0.0;
    if ((rc = trial(eps, &rp))) return rc;
    if ((rc = qs::dot(c, lhs, rs, &SR))) return rc;
    if ((rc = qs::dot(c, lhs, S->out, &SPR))) return rc;
    const double ta = -1.0 * eps * SR / (SPR - SR);
    if (ta > -0.1 && ta < 10.0) cand.push_back(ta);
    for (double a : {0.1, 0.2, 0.3, 0.5, 0.75, 1.0}) cand.push_back(a);
    for (double a : cand) {
      double r = 0.0;
      if ((rc = trial(a, &r))) return rc;
      if (std::isfinite(r) && r < bestres) { best = a; bestres = r; }
    }
    if ((r0 - bestres) / r0 > 0.02) { *alpha_out = best; return PDG_OK; }
    // brute force
    cand.clear();
    for (int i = 0; i < 25; ++i) cand.push_back(0.002 * (i + 1));
    for (int i = 0; i < 15; ++i) cand.push_back(0.01 * (i + 1) + 0.05);
    for (int i = 0; i < 6; ++i) cand.push_back(0.05 * (i + 1) + 0.2);
    for (double a : {0.75, 1.0, 1.5, 2.0, 5.0, 10.0}) cand.push_back(a);
    std::vector<std::pair<double, double>> up;  // (alpha, relative increase)
    for (double a : cand) {
      double r = 0.0;
      if ((rc = trial(a, &r))) return rc;
      if (std::isfinite(r)) {
        if (r < bestres) { best = a; bestres = r; }
        else if (r > r0) up.push_back({a, (r - r0) / r0});
      }
    }
    if ((r0 - bestres) / r0 > 0.0001) { *alpha_out = best; return PDG_OK; }
    // no reduction: the alpha whose increase is closest to 10%
    double bestdev = 1.0e50;
    for (const auto& pr : up) {
      if (std::fabs(pr.second - 0.1) < bestdev) { best = pr.first; bestdev = std::fabs(pr.second - 0.1); }
    }
    *alpha_out = best;
    return PDG_OK;
  }

  // damped-Newton diagonal (quasiStaticsDampTangent, :3302-3319) from the
  // diagonal of K at the current state: d' = f d + g ||f d||_inf
  int damp(double scale, double shift) {
    double n2 = 0.0, ninf = 0.0;
    int rc = qs::norms(c, S->diag, nullptr, &n2, &ninf);
    if (rc) return rc;
    hipLaunchKernelGGL(k_damp_vec, dim3(vb), dim3(RB), 0, c->stream, n3, S->diag, scale,
                       shift * (scale * ninf), S->dampv);
    S->damped = true;
    return c->hip(hipGetLastError(), "damp");
  }
};

}  // namespace
}  // namespace pdg

extern "C" {

int pdg_qs_params_default(pdg_qs_params* p) {
  if (!p) return PDG_E_ARG;
  p->maxSolverIterations = 10;
  p->tolerance = 1.0e-6;
  p->absoluteTolerance = 0;
  p->disableHeuristics = 0;
  p->dampedNewtonDiagonalScaleFactor = 1.0001;
  p->dampedNewtonDiagonalShiftFactor = 0.00001;
  p->adaptiveLoadStepping = 0;
  p->maxLoadStepReductionsInOneStep = 0;
  p->maxTotalLoadStepReductions = 0;
  p->reduceAllRemainingLoadSteps = 0;
  p->precond = PDG_PRECOND_NONE;
  p->linearTolerance = 1.0e-4;
  p->maxLinearIterations = 0;
  return PDG_OK;
}

int pdg_qs_run(pdg_ctx* h, const pdg_bc* bcs, int numBcs, const double* times, int numTimes,
               const pdg_qs_params* prm, double* u, pdg_qs_step* steps, int maxSteps,
               int* numSteps) {
  return pdg_qs_run_record(h, bcs, numBcs, times, numTimes, prm, u, steps, maxSteps, numSteps,
                           nullptr);
}

int pdg_qs_run_record(pdg_ctx* h, const pdg_bc* bcs, int numBcs, const double* times,
                      int numTimes, const pdg_qs_params* prm, double* u, pdg_qs_step* steps,
                      int maxSteps, int* numSteps, double* uSteps) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !u || !prm || !times || numTimes < 1 || numBcs < 0 ||
      (numBcs > 0 && !bcs) || prm->precond < 0 || prm->precond > 3)
    return PDG_E_ARG;
  if (prm->adaptiveLoadStepping &&
      prm->maxTotalLoadStepReductions < prm->maxLoadStepReductionsInOneStep)
    return c->fail(PDG_E_ARG, "'Maximum total load step reductions' cannot be smaller than "
                              "'Maximum load step reductions in one step'");
  for (int k = 1; k < numTimes; k++)
    if (!(times[k] > times[k - 1])) return c->fail(PDG_E_ARG, "load-step times must increase");
  (void)hipSetDevice(c->device);
  int rc = qs::ensure(c);
  if (rc) return rc;
  QsState* S = c->qs;
  S->coarse_ok = false;  // the two-level preconditioner is rebuilt for the run's DOFs
  const int N = c->N, n3 = 3 * N;
  hipStream_t s = c->stream;
  // kinematic DOFs (internal order) with their u(t) = c1 t + c0
  std::vector<int> kidx(n3, -1), kdof;
  std::vector<double> kc0, kc1;
  for (int b = 0; b < numBcs; b++) {
    if (bcs[b].coord < 0 || bcs[b].coord > 2 || bcs[b].numNodes < 0 || (bcs[b].numNodes && !bcs[b].nodes))
      return c->fail(PDG_E_ARG, "invalid boundary condition");
    for (int q = 0; q < bcs[b].numNodes; q++) {
      const int p = bcs[b].nodes[q];
      if (p < 0 || p >= N) return c->fail(PDG_E_ARG, "boundary-condition node out of range");
      const int d = 3 * c->iperm[p] + bcs[b].coord;
      if (kidx[d] < 0) {
        kidx[d] = (int)kdof.size();
        kdof.push_back(d);
        kc0.push_back(0.0);
        kc1.push_back(0.0);
      }
      kc0[kidx[d]] = bcs[b].c0;
      kc1[kidx[d]] = bcs[b].c1;
    }
  }
  S->hfixed.assign(n3, 0);
  for (int d : kdof) S->hfixed[d] = 1;
  rc = c->hip(hipMemcpyAsync(S->fixed, S->hfixed.data(), n3, hipMemcpyHostToDevice, s), "h2d fixed");
  for (int k = 0; k < 5 && !rc; k++)
    if (!S->rv[k]) rc = c->hip(hipMalloc((void**)&S->rv[k], std::max<size_t>(8 * (size_t)n3, 64)), "qs alloc");
  if (rc) return rc;
  QsRun R{c, S, n3, (unsigned)((n3 + RB - 1) / RB), S->rv[0], S->rv[1], S->rv[2], S->rv[3], S->rv[4], S->xs};
  if ((rc = qs::to_dev(c, u, R.ub))) return rc;
  rc = c->hip(hipMemsetAsync(R.pred, 0, 8 * (size_t)n3, s), "zero predictor");
  std::vector<double> hub(n3), hdu(n3);
  if (!rc) rc = qs::to_host(c, R.ub, u);  // reference order, for the zero-increment rule
  if (rc) return rc;
  for (int k = 0; k < N; k++)
    for (int q = 0; q < 3; q++) hub[3 * (size_t)k + q] = u[3 * (size_t)c->perm[k] + q];
  const bool adaptive = prm->adaptiveLoadStepping != 0, noheur = prm->disableHeuristics != 0;
  const int maxLin = prm->maxLinearIterations > 0 ? prm->maxLinearIterations : n3;
  std::vector<double> ts(times, times + numTimes);
  const double tfinal = ts.back();
  double tc = ts[0], tp = tc;
  bool failed = false, failed_in_step = false;
  int fail_step = 0, fail_total = 0, nrec = 0;
  int result = PDG_OK;
  for (int step = 1; step < (int)ts.size(); step++) {
    if (!adaptive || !failed) tp = tc;
    failed = false;
    tc = ts[step];
    const double dt = tc - tp;
    // the temperature change at the load step's time (prescribed
    // temperatures are applied with the displacements)
    if (c->thermal()) {
      c->temp_time = tc;
      if ((rc = qs::thermal_sync(c))) return rc;
    }
    // deltaU: the prescribed increments (DirichletIncrementBC with coefficient
    // 1, src/core/Peridigm_BoundaryCondition.cpp:202-241; a zero increment
    // takes the field's own value as the previous one, :136-143)
    std::fill(hdu.begin(), hdu.end(), 0.0);
    if (tc != tp) {
      for (size_t k = 0; k < kdof.size(); k++) {
        const double cur = kc1[k] * tc + kc0[k];
        double prev = kc1[k] * tp + kc0[k];
        if (cur - prev == 0.0) prev = hub[kdof[k]];
        hdu[kdof[k]] = 1.0 * (cur - prev) + 0.0 * (cur - prev) * (1.0 / dt);
      }
    }
    rc = c->hip(hipMemcpyAsync(R.du, hdu.data(), 8 * (size_t)n3, hipMemcpyHostToDevice, s), "h2d du");
    if (rc) return rc;
    R.state(0.0);
    S->damped = false;
    double mult = 1.0, rnorm = 0.0;
    if (!prm->absoluteTolerance && (rc = qs::reaction_norm(c, &mult))) return rc;
    if ((rc = qs::residual_norm(c, &rnorm))) return rc;
    c->qs_tol_multiplier = mult;
    const double tol = prm->tolerance * mult;
    int it = 1, lits = 0;
    bool damped = false;
    int prec = prm->precond;
    double alpha = 0.0;
    while (rnorm > tol && it <= prm->maxSolverIterations) {
      bool conv = true;
      if (it == 1 && step > 1 && !noheur) {
        // predictor: the previous load step's velocity times this increment
        hipLaunchKernelGGL(k_lin2, dim3(R.vb), dim3(RB), 0, s, n3, dt, R.pred, 0.0, R.pred, R.lhs);
        hipLaunchKernelGGL(k_zero_fixed, dim3(R.vb), dim3(RB), 0, s, n3, S->fixed, R.lhs);
      } else {
        if (it > 50 && !damped && !noheur) damped = true;
        // tangent at the current state: diagonal blocks (preconditioner and
        // damping), then the linear solve K lhs = r
        if ((rc = qs::block_jacobi(c))) return rc;
        if (prec == 3 && !S->coarse_ok && (rc = qs::coarse_build(c))) return rc;
        if (damped && (rc = R.damp(prm->dampedNewtonDiagonalScaleFactor,
                                   prm->dampedNewtonDiagonalShiftFactor)))
          return rc;
        if ((rc = qs::residual(c))) return rc;
        auto solve = [&]() -> int {
          int r2 = c->hip(hipMemcpyAsync(S->b, S->out, 8 * (size_t)n3, hipMemcpyDeviceToDevice, s), "copy");
          if (!r2) r2 = c->hip(hipMemsetAsync(R.lhs, 0, 8 * (size_t)n3, s), "zero");
          if (r2) return r2;
          int li = 0;
          double rel = 0.0;
          r2 = qs_solve(c, R.lhs, S->b, prec, prm->linearTolerance, maxLin, &li, &rel);
          lits += li;
          return r2;
        };
        rc = solve();
        if (rc == PDG_E_NOCONV) {
          conv = false;
          if (!noheur) {
            // damp the tangent, drop the preconditioner and try again
            if (!damped && (rc = R.damp(prm->dampedNewtonDiagonalScaleFactor,
                                        prm->dampedNewtonDiagonalShiftFactor)))
              return rc;
            damped = true;
            prec = PDG_PRECOND_NONE;
            rc = solve();
            conv = rc == PDG_OK;
            if (rc == PDG_E_NOCONV) rc = PDG_OK;
          } else {
            rc = PDG_OK;
          }
        }
        if (rc) return rc;
      }
      if (!conv) { rnorm = 1.0e50; break; }
      hipLaunchKernelGGL(k_zero_fixed, dim3(R.vb), dim3(RB), 0, s, n3, S->fixed, R.lhs);
      alpha = 1.0;
      if (!noheur && (rc = R.line_search(&alpha))) return rc;
      hipLaunchKernelGGL(k_lin2, dim3(R.vb), dim3(RB), 0, s, n3, 1.0, R.du, alpha, R.lhs, R.du);
      R.state(0.0);
      if ((rc = qs::residual_norm(c, &rnorm))) return rc;
      it++;
    }
    S->damped = false;
    int status = 0;
    if (rnorm > tol) {
      if (rnorm < 100.0 * tol) {
        status = 1;  // accepted with a warning
      } else if (adaptive) {
        fail_step++;
        fail_total++;
        if (fail_total > prm->maxTotalLoadStepReductions ||
            fail_step > prm->maxLoadStepReductionsInOneStep) {
          result = c->fail(PDG_E_NOCONV, "quasi-static solve failed: load-step reductions exhausted");
          break;
        }
        failed = true;
        failed_in_step = true;
      } else {
        result = c->fail(PDG_E_NOCONV, "quasi-static solve failed to converge");
        break;
      }
    }
    if (adaptive && failed) {
      ts.insert(ts.begin() + step, tp + dt / 2.0);
      step--;
    }
    if (adaptive && prm->reduceAllRemainingLoadSteps && !failed && failed_in_step) {
      ts.erase(ts.begin() + step + 1, ts.end());
      for (int i = 0; ts[step + i] < tfinal; i++) ts.push_back(tc + (i + 1) * dt);
    }
    if (!adaptive || !failed) {
      failed_in_step = false;
      fail_step = 0;
      // u += deltaU; predictor = deltaU / dt
      hipLaunchKernelGGL(k_div, dim3(R.vb), dim3(RB), 0, s, n3, R.du, dt, R.pred);
      hipLaunchKernelGGL(k_lin2, dim3(R.vb), dim3(RB), 0, s, n3, 1.0, R.ub, 1.0, R.du, R.ub);
      rc = c->hip(hipMemcpyAsync(hub.data(), R.ub, 8 * (size_t)n3, hipMemcpyDeviceToHost, s), "d2h");
      if (!rc) rc = c->hip(hipStreamSynchronize(s), "sync");
      if (rc) return rc;
      if (uSteps && nrec < maxSteps)
        for (int k = 0; k < N; k++)
          for (int q = 0; q < 3; q++)
            uSteps[(size_t)nrec * n3 + 3 * (size_t)c->perm[k] + q] = hub[3 * (size_t)k + q];
      if (steps && nrec < maxSteps) {
        pdg_qs_step& st = steps[nrec];
        st.timePrevious = tp;
        st.timeCurrent = tc;
        st.newtonIterations = it - 1;
        st.linearIterations = lits;
        st.residualNorm = rnorm;
        st.toleranceMultiplier = mult;
        st.status = status;
        st.dampedNewton = damped ? 1 : 0;
        st.lastAlpha = alpha;
      }
      nrec++;
    }
  }
  S->damped = false;
  if (numSteps) *numSteps = nrec;
  rc = qs::to_host(c, R.ub, u);
  if (rc) return rc;
  return result;
}

}  // extern "C"
