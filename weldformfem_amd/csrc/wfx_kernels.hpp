// HIP kernels of the per-step Domain_d API, one kernel per reference
// function, templated on the element real type T (double: FP64 path, float:
// FP32-mixed path).  Nodal kinematics (x, u, u_dt, v, a) stay double in both.
//
// This header is included by exactly two translation units:
//   wfx_k64.hip  instantiates T=double, compiled with -ffp-contract=off so
//                that every expression rounds like the reference's CPU code;
//   wfx_k32.hip  instantiates T=float with contraction allowed.
#pragma once

#include <cfloat>
#include <type_traits>

#include "wfx_internal.hpp"

namespace wfx {

constexpr int kBlock = 256;
inline int nblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

// Kernels have internal linkage: each precision translation unit carries its
// own copy of the precision-independent ones.
namespace {

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), and each XCD has its own L2.  Remap so that every XCD walks one
// contiguous range of logical blocks: neighbouring elements share nodes, so
// their gathers then hit the same L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int per = nb >> 3, rem = nb & 7, xcd = b & 7, idx = b >> 3;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}
__device__ __forceinline__ int gtid() { return xcd_block() * blockDim.x + threadIdx.x; }

// ---------------------------------------------------------------------------
// Per-step CFL time step (Solver_explicit.C:576-600, the BUILD_REMESH CPU
// solver): dt = cfl * min edge / (sqrt(K / rho[0]) + max |v|), from the
// positions and the predicted velocities of the step.  The passes that
// touch the elements and the nodes leave one partial minimum (edge length)
// or maximum (speed) per block; one small kernel reduces them and writes dt
// to the context's block of scalars ([kCflDt], with the extremes and the
// sound speed beside it), where the later kernels of the step and the next
// prediction read it.  NaNs are skipped, as the reference's `<` and `>`
// comparisons skip them.
constexpr double kMinEdgeInit = 1.0e6;  // calcMinEdgeLength's start value (Domain_d.C:2225)

// the block's minimum (MIN) or maximum of v, valid in thread 0
template <int BS, bool MIN>
__device__ __forceinline__ double block_extreme(double v) {
  __shared__ double sh[BS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o);
    v = MIN ? fmin(v, w) : fmax(v, w);
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < BS / 64; ++k) r = MIN ? fmin(r, sh[k]) : fmax(r, sh[k]);
  return r;
}

// smallest edge of the tet of an element's first four nodes
// (calcMinEdgeLength, Domain_d.C:2224-2260)
template <typename T>
__device__ __forceinline__ double min_edge4(const T (&P)[4][3]) {
  const int ep[6][2] = {{1, 0}, {2, 0}, {3, 0}, {2, 1}, {3, 1}, {3, 2}};
  double m = kMinEdgeInit;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const T ex = P[ep[i][0]][0] - P[ep[i][1]][0], ey = P[ep[i][0]][1] - P[ep[i][1]][1],
            ez = P[ep[i][0]][2] - P[ep[i][1]][2];
    const double len = (double)sqrt(ex * ex + ey * ey + ez * ez);
    if (len < m) m = len;
  }
  return m;
}

__global__ __launch_bounds__(kBlock) void k_cfl_min_edge(int E, int N, const int *__restrict__ elnod,
                                                         const double *__restrict__ x, double *__restrict__ part) {
  const int e = gtid();
  double m = kMinEdgeInit;
  if (e < E) {
    double P[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int g = elnod[(long)k * E + e];
#pragma unroll
      for (int c = 0; c < 3; ++c) P[k][c] = x[(long)c * N + g];
    }
    m = min_edge4<double>(P);
  }
  m = block_extreme<kBlock, true>(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// the largest nodal speed |v| (norm(getVelVec(i)))
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cfl_max_speed(int N, const T *__restrict__ v,
                                                          double *__restrict__ part) {
  const int i = gtid();
  double sp = 0.0;
  if (i < N) {
    const double vx = (double)v[i], vy = (double)v[(long)N + i], vz = (double)v[2L * N + i];
    sp = sqrt(vx * vx + vy * vy + vz * vz);
  }
  sp = block_extreme<kBlock, false>(sp);
  if (threadIdx.x == 0) part[blockIdx.x] = sp;
}

// dt of a step from the partial extremes (one block)
constexpr int kCflReduceBlock = 1024;
template <typename T>
__global__ __launch_bounds__(kCflReduceBlock) void k_cfl_reduce(
    const double *__restrict__ minl_part, int nmin, const double *__restrict__ vmax_part, int nmax,
    const T *__restrict__ rho, double K, double cfl, double *__restrict__ s, double *__restrict__ dt_hist,
    int step) {
  double m = kMinEdgeInit, v = 0.0;
  for (int k = threadIdx.x; k < nmin; k += kCflReduceBlock) m = fmin(m, minl_part[k]);
  for (int k = threadIdx.x; k < nmax; k += kCflReduceBlock) v = fmax(v, vmax_part[k]);
  m = block_extreme<kCflReduceBlock, true>(m);
  v = block_extreme<kCflReduceBlock, false>(v);
  if (threadIdx.x == 0) {
    const double cs = sqrt(K / (double)rho[0]);
    const double dt = cfl * m / (cs + v);
    s[kCflMinl] = m;
    s[kCflVmax] = v;
    s[kCflCs] = cs;
    s[kCflDt] = dt;
    if (dt_hist) dt_hist[step] = dt;
  }
}

// ---------------------------------------------------------------------------
// tensor3 with the reference's operators (include/common/Tensor3.C)
// ---------------------------------------------------------------------------
template <typename T>
struct T3 {
  T xx, xy, xz, yx, yy, yz, zx, zy, zz;
};
template <typename T>
__device__ __forceinline__ T3<T> t3_from_sym(T a0, T a1, T a2, T a3, T a4, T a5) {
  T3<T> r;
  r.xx = a0; r.yy = a1; r.zz = a2;
  r.xy = r.yx = a3; r.yz = r.zy = a4; r.xz = r.zx = a5;
  return r;
}
// FromFlatAntiSym with a zero diagonal (the reference stores zeros there)
template <typename T>
__device__ __forceinline__ T3<T> t3_from_antisym(T w01, T w12, T w02) {
  T3<T> r;
  r.xx = T(0); r.yy = T(0); r.zz = T(0);
  r.xy = w01; r.yz = w12; r.xz = w02;
  r.yx = -w01; r.zy = -w12; r.zx = -w02;
  return r;
}
template <typename T>
__device__ __forceinline__ T t3_trace(const T3<T> &m) { return m.xx + m.yy + m.zz; }
template <typename T>
__device__ __forceinline__ T3<T> t3_identity() {
  T3<T> r;
  r.xx = r.yy = r.zz = T(1);
  r.xy = r.xz = r.yx = r.yz = r.zx = r.zy = T(0);
  return r;
}
// Tensor3.C operator*(tensor3,tensor3): x column = A*B, y/z columns = A*B^T.
template <typename T>
__device__ __forceinline__ T3<T> t3_mul(const T3<T> &a, const T3<T> &b) {
  T3<T> r;
  r.xx = a.xx * b.xx + a.xy * b.yx + a.xz * b.zx;
  r.xy = a.xx * b.yx + a.xy * b.yy + a.xz * b.yz;
  r.xz = a.xx * b.zx + a.xy * b.zy + a.xz * b.zz;
  r.yx = a.yx * b.xx + a.yy * b.yx + a.yz * b.zx;
  r.yy = a.yx * b.yx + a.yy * b.yy + a.yz * b.yz;
  r.yz = a.yx * b.zx + a.yy * b.zy + a.yz * b.zz;
  r.zx = a.zx * b.xx + a.zy * b.yx + a.zz * b.zx;
  r.zy = a.zx * b.yx + a.zy * b.yy + a.zz * b.yz;
  r.zz = a.zx * b.zx + a.zy * b.zy + a.zz * b.zz;
  return r;
}
template <typename T>
__device__ __forceinline__ T3<T> t3_trans(const T3<T> &m) {
  T3<T> r;
  r.xx = m.xx; r.yy = m.yy; r.zz = m.zz;
  r.xy = m.yx; r.xz = m.zx; r.yx = m.xy; r.yz = m.zy; r.zx = m.xz; r.zy = m.yz;
  return r;
}
template <typename T>
__device__ __forceinline__ T3<T> t3_scale(const T3<T> &m, T f) {
  T3<T> r;
  r.xx = m.xx * f; r.xy = m.xy * f; r.xz = m.xz * f;
  r.yx = m.yx * f; r.yy = m.yy * f; r.yz = m.yz * f;
  r.zx = m.zx * f; r.zy = m.zy * f; r.zz = m.zz * f;
  return r;
}
template <typename T>
__device__ __forceinline__ T3<T> t3_add(const T3<T> &a, const T3<T> &b) {
  T3<T> r;
  r.xx = a.xx + b.xx; r.xy = a.xy + b.xy; r.xz = a.xz + b.xz;
  r.yx = a.yx + b.yx; r.yy = a.yy + b.yy; r.yz = a.yz + b.yz;
  r.zx = a.zx + b.zx; r.zy = a.zy + b.zy; r.zz = a.zz + b.zz;
  return r;
}
template <typename T>
__device__ __forceinline__ T3<T> t3_sub(const T3<T> &a, const T3<T> &b) {
  T3<T> r;
  r.xx = a.xx - b.xx; r.xy = a.xy - b.xy; r.xz = a.xz - b.xz;
  r.yx = a.yx - b.yx; r.yy = a.yy - b.yy; r.yz = a.yz - b.yz;
  r.zx = a.zx - b.zx; r.zy = a.zy - b.zy; r.zz = a.zz - b.zz;
  return r;
}

// ---------------------------------------------------------------------------
// nodal update kernels: flat over the [3][N] component-major arrays
// ---------------------------------------------------------------------------

// UpdatePrediction (Domain_d.C:961-974); with dtp the step length is read
// from device memory (the per-step CFL time step of the previous step)
__global__ __launch_bounds__(kBlock) void k_update_prediction(long n3, double *__restrict__ v,
                                                              const double *__restrict__ prev_a,
                                                              double *__restrict__ u_dt, double dt,
                                                              double beta, double gamma,
                                                              const double *__restrict__ dtp) {
  long i = gtid();
  if (i >= n3) return;
  if (dtp) dt = *dtp;
  double pa = prev_a[i], vi = v[i];
  u_dt[i] = dt * (vi + (0.5 - beta) * dt * pa);
  v[i] = vi + (1.0 - gamma) * dt * pa;
}

// ImposeBCV / ImposeBCA (Domain_d.C:1109-1134) on one component array.
// Lists are de-duplicated on the host (last entry per node wins, as in the
// reference's serial loop), so no two threads touch the same address.
__global__ __launch_bounds__(kBlock) void k_impose_value(int count, const int *__restrict__ nod,
                                                         const double *__restrict__ val,
                                                         double *__restrict__ comp) {
  int i = gtid();
  if (i >= count) return;
  comp[nod[i]] = val ? val[i] : 0.0;
}

// calcAccel (Mechanical.C:321-341)
__global__ __launch_bounds__(kBlock) void k_calc_accel(int N, const double *__restrict__ fe,
                                                       const double *__restrict__ fi,
                                                       const double *__restrict__ m,
                                                       double *__restrict__ a,
                                                       const double *__restrict__ cf) {
  int i = gtid();
  if (i >= N) return;
  double mi = m[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double ad = (fe[(long)d * N + i] - fi[(long)d * N + i]) / mi;
    if (cf) ad += cf[(long)d * N + i] / mi;  // contact force (Mechanical.C:330-335)
    a[(long)d * N + i] = ad;
  }
}

// UpdateCorrectionAccVel (Domain_d.C:981-997)
__global__ __launch_bounds__(kBlock) void k_correction_accvel(long n3, double *__restrict__ a,
                                                              const double *__restrict__ prev_a,
                                                              double *__restrict__ v, double f,
                                                              double alpha, double gdt) {
  long i = gtid();
  if (i >= n3) return;
  double ai = f * (a[i] - alpha * prev_a[i]);
  a[i] = ai;
  v[i] = v[i] + gdt * ai;
}

// UpdateCorrectionPos (Domain_d.C:1005-1025); bdt2 = beta*dt*dt
__global__ __launch_bounds__(kBlock) void k_correction_pos(long n3, const double *__restrict__ a,
                                                           double *__restrict__ u_dt,
                                                           double *__restrict__ x,
                                                           double *__restrict__ prev_a,
                                                           double *__restrict__ u, double bdt2) {
  long i = gtid();
  if (i >= n3) return;
  double ai = a[i];
  double ud = u_dt[i] + bdt2 * ai;
  u_dt[i] = ud;
  x[i] = x[i] + ud;
  prev_a[i] = ai;
  u[i] = u[i] + ud;
}

// ---------------------------------------------------------------------------
// element kernels
// ---------------------------------------------------------------------------

// Inverted or degenerate element: detJ <= 0 or not finite (the check of
// Solver_explicit.C:639-652, which prints and counts them).  Counted in the
// context's status block; the step goes on as the reference's does.
template <typename T>
__device__ __forceinline__ void count_inverted(T det, unsigned int *status) {
  if (status && !(det > T(0))) atomicAdd(status + kStatInverted, 1u);
}

// calcElemJAndDerivatives (Domain_d.C:1701-2044): hexa :1822-1850, tet
// :1855-1887, determinant Matrix.h:593-618, adjugate Matrix.h:694-727.
template <typename T, int NN>
__global__ __launch_bounds__(kBlock) void k_elem_jacobians(int E, int N, const int *__restrict__ elnod,
                                                           const double *__restrict__ x,
                                                           T *__restrict__ detJ, T *__restrict__ dH,
                                                           unsigned int *__restrict__ status) {
  int e = gtid();
  if (e >= E) return;
  T X[NN][3];
  if constexpr (std::is_same<T, double>::value) {
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      int g = elnod[(long)a * E + e];
#pragma unroll
      for (int c = 0; c < 3; ++c) X[a][c] = x[(long)c * N + g];
    }
  } else {
    // FP32 bulk: coordinates relative to local node 0, differenced in FP64
    int g0 = elnod[e];
    double x0[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) x0[c] = x[(long)c * N + g0];
#pragma unroll
    for (int c = 0; c < 3; ++c) X[0][c] = T(0);
#pragma unroll
    for (int a = 1; a < NN; ++a) {
      int g = elnod[(long)a * E + e];
#pragma unroll
      for (int c = 0; c < 3; ++c) X[a][c] = T(x[(long)c * N + g] - x0[c]);
    }
  }
  T J[3][3];
  if constexpr (NN == 8) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      J[0][c] = T(0.125) * (-X[0][c] + X[1][c] + X[2][c] - X[3][c] - X[4][c] + X[5][c] + X[6][c] - X[7][c]);
      J[1][c] = T(0.125) * (-X[0][c] - X[1][c] + X[2][c] + X[3][c] - X[4][c] - X[5][c] + X[6][c] + X[7][c]);
      J[2][c] = T(0.125) * (-X[0][c] - X[1][c] - X[2][c] - X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      J[0][c] = X[1][c] - X[0][c];
      J[1][c] = X[2][c] - X[0][c];
      J[2][c] = X[3][c] - X[0][c];
    }
  }
  T C[3][3];
  C[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
  C[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
  C[0][2] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  C[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
  C[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
  C[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
  C[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
  C[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
  C[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
  const T det = J[0][0] * J[1][1] * J[2][2] - J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2] +
                J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] - J[0][2] * J[1][1] * J[2][0];
  detJ[e] = det;
  count_inverted(det, status);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // adj(d,i) = C(i,d)
    T a0 = C[0][d], a1 = C[1][d], a2 = C[2][d];
    T *o = dH + (long)d * NN * E + e;
    if constexpr (NN == 8) {
      const T q = T(0.125);
      o[0 * (long)E] = q * (-a0 - a1 - a2);
      o[1 * (long)E] = q * (a0 - a1 - a2);
      o[2 * (long)E] = q * (a0 + a1 - a2);
      o[3 * (long)E] = q * (-a0 + a1 - a2);
      o[4 * (long)E] = q * (-a0 - a1 + a2);
      o[5 * (long)E] = q * (a0 - a1 + a2);
      o[6 * (long)E] = q * (a0 + a1 + a2);
      o[7 * (long)E] = q * (-a0 + a1 + a2);
    } else {
      o[0 * (long)E] = -a0 - a1 - a2;
      o[1 * (long)E] = a0;
      o[2 * (long)E] = a1;
      o[3 * (long)E] = a2;
    }
  }
}

// CalcElemVol (Mechanical.C:264-293): vol = detJ*w*1.0, w = 8 or 1/6
template <typename T>
__global__ __launch_bounds__(kBlock) void k_elem_vol(int E, const T *__restrict__ detJ,
                                                     T *__restrict__ vol, T w) {
  int e = gtid();
  if (e >= E) return;
  T v = T(0);
  v += detJ[e] * w * T(1);
  vol[e] = v;
}

// calcElemDensity (Mechanical.C:295-307)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_elem_density(int E, const T *__restrict__ rho0,
                                                         const T *__restrict__ vol0,
                                                         const T *__restrict__ vol,
                                                         T *__restrict__ rho) {
  int e = gtid();
  if (e >= E) return;
  rho[e] = rho0[e] * vol0[e] / vol[e];
}

// Walk of node i's CSR entries [o, o + c) in batches of kCsrBatch, over the
// packed entries (element << 3) | local node: the first eight come from the
// padded table pad[j * N + i] (unit stride across a wavefront, one or two
// cache lines per load instead of sixteen for a strided CSR read), the rest
// from the packed CSR.  The batch's entries, then its data loads, are all
// issued (positions past the node's count repeat its last entry: a load
// behind a branch is waited for on its own), and acc(value) runs over the
// node's own entries in CSR order (sums as in a plain loop, bit for bit).
constexpr int kCsrBatch = 8;
template <class Load, class Acc>
__device__ __forceinline__ void csr_walk(int i, int N, int o, int c, const int *__restrict__ pad,
                                         const int *__restrict__ pk, Load load, Acc acc) {
  using V = decltype(load(0, 0));
  for (int k = 0; k < c; k += kCsrBatch) {
    int ent[kCsrBatch];
#pragma unroll
    for (int j = 0; j < kCsrBatch; ++j) {
      const int q = k + j < c ? k + j : c - 1;
      const int *src = q < 8 ? pad + (long)q * N + i : pk + o + q;
      ent[j] = *src;
    }
    V val[kCsrBatch];
#pragma unroll
    for (int j = 0; j < kCsrBatch; ++j) val[j] = load(ent[j] >> 3, ent[j] & 7);
#pragma unroll
    for (int j = 0; j < kCsrBatch; ++j)
      if (k + j < c) acc(val[j]);
  }
}

// CalcNodalVol (Mechanical.C:1555-1572): deterministic CSR gather
template <typename T>
__global__ __launch_bounds__(kBlock) void k_nodal_vol(int N, const int *__restrict__ off,
                                                      const int *__restrict__ cnt,
                                                      const int *__restrict__ pad,
                                                      const int *__restrict__ pk,
                                                      const T *__restrict__ vol,
                                                      double *__restrict__ voln, int nodxelem) {
  int i = gtid();
  if (i >= N) return;
  int o = off[i], c = cnt[i];
  double s = 0.0;
  csr_walk(i, N, o, c, pad, pk, [&](int e, int) { return (double)vol[e]; }, [&](double x) { s += x; });
  s /= nodxelem;
  voln[i] = s;
}

// CalcNodalMassFromVol (Mechanical.C:1576-1601)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_nodal_mass(int N, const int *__restrict__ off,
                                                       const int *__restrict__ cnt,
                                                       const int *__restrict__ pad,
                                                       const int *__restrict__ pk,
                                                       const T *__restrict__ rho,
                                                       const double *__restrict__ voln,
                                                       double *__restrict__ mdiag) {
  int i = gtid();
  if (i >= N) return;
  int o = off[i], c = cnt[i];
  double vn = voln[i], dc = (double)c, mass = 0.0;
  csr_walk(i, N, o, c, pad, pk, [&](int e, int) { return (double)rho[e]; },
           [&](double r) { mass += 1.0 * r * vn / dc; });
  mdiag[i] = mass;
}

// calcElemStrainRates (Mechanical.C:41-126)
template <typename T, int NN>
__global__ __launch_bounds__(kBlock) void k_strain_rates(int E, int N, const int *__restrict__ elnod,
                                                         const double *__restrict__ v,
                                                         const T *__restrict__ dH,
                                                         const T *__restrict__ detJ,
                                                         T *__restrict__ sr, T *__restrict__ rr) {
  int e = gtid();
  if (e >= E) return;
  T s00 = 0, s11 = 0, s22 = 0, s01 = 0, s12 = 0, s02 = 0, w01 = 0, w12 = 0, w02 = 0;
  T f = T(1) / detJ[e];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    int g = elnod[(long)a * E + e];
    T vx = (T)v[g], vy = (T)v[(long)N + g], vz = (T)v[2L * N + g];
    T h0 = dH[(long)(0 * NN + a) * E + e], h1 = dH[(long)(1 * NN + a) * E + e],
      h2 = dH[(long)(2 * NN + a) * E + e];
    s00 = s00 + h0 * f * vx;
    s11 = s11 + h1 * f * vy;
    s22 = s22 + h2 * f * vz;
    s01 = s01 + f * (h1 * vx + h0 * vy);
    w01 = w01 + f * (h1 * vx - h0 * vy);
    s12 = s12 + f * (h2 * vy + h1 * vz);
    s02 = s02 + f * (h2 * vx + h0 * vz);
    w12 = w12 + f * (h2 * vy - h1 * vz);
    w02 = w02 + f * (h2 * vx - h0 * vz);
  }
  const long Es = E;
  sr[0 * Es + e] = s00; sr[1 * Es + e] = s11; sr[2 * Es + e] = s22;
  sr[3 * Es + e] = s01 * T(0.5); sr[4 * Es + e] = s12 * T(0.5); sr[5 * Es + e] = s02 * T(0.5);
  rr[0 * Es + e] = w01 * T(0.5); rr[1 * Es + e] = w12 * T(0.5); rr[2 * Es + e] = w02 * T(0.5);
}

// calcElemPressure node part (Mechanical.C:698-705): CSR sums of vol, vol0
// and their ratio, the value the element part averages (:720-730)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pressure_nodes(int N, const int *__restrict__ off,
                                                           const int *__restrict__ cnt,
                                                           const int *__restrict__ pad,
                                                           const int *__restrict__ pk,
                                                           const T *__restrict__ vol,
                                                           const T *__restrict__ vol0,
                                                           double *__restrict__ ratio) {
  int i = gtid();
  if (i >= N) return;
  int o = off[i], c = cnt[i];
  double s0 = 0.0, s = 0.0;
  csr_walk(i, N, o, c, pad, pk, [&](int e, int) { return double2{(double)vol0[e], (double)vol[e]}; },
           [&](double2 q) {
             s0 += q.x;
             s += q.y;
           });
  ratio[i] = s / s0;
}

struct PressParams {
  double K, alpha_free, log_factor, J_min, hg_coeff_free, pspg_scale, p_pspg_bulkfac,
      av_coeff_div, av_coeff_bulk;
  int full;  // 0: all stabilisation terms that need h, c, div v are zero
  double alpha_contact, hg_coeff_contact;
  // contact forces [3][N] when an element's contact state can change its
  // pressure (contact on and contact coefficients that differ from the free
  // ones, or the terms of `full`); null otherwise
  const double *cf;
  int N;
  // fused path: 1 when an element's initial volume or density may make a
  // term non-finite (found by the host scan of vol0 and rho), so that every
  // element evaluates the pressure and viscosity terms in full
  int guard;
};

// Parameters of calcElemPressure for the context (per-step and fused paths).
inline PressParams press_params(const wfx_ctx *c) {
  const auto &st = c->prm.stab;
  PressParams pp;
  pp.K = c->prm.mat.K;
  pp.alpha_free = st.alpha_free;
  pp.log_factor = st.log_factor;
  pp.J_min = st.J_min;
  pp.hg_coeff_free = st.hg_coeff_free;
  pp.pspg_scale = st.pspg_scale;
  pp.p_pspg_bulkfac = st.p_pspg_bulkfac;
  pp.av_coeff_div = st.av_coeff_div;
  pp.av_coeff_bulk = st.av_coeff_bulk;
  pp.full = (st.pspg_scale != 0.0 || st.p_pspg_bulkfac != 0.0 || st.av_coeff_div != 0.0 ||
             st.av_coeff_bulk != 0.0);
  pp.alpha_contact = st.alpha_contact;
  pp.hg_coeff_contact = st.hg_coeff_contact;
  const bool blend = pp.full || st.alpha_contact != st.alpha_free || st.hg_coeff_contact != st.hg_coeff_free;
  pp.cf = (c->contact.on && blend) ? c->contact.contforce : nullptr;
  pp.N = c->N;
  pp.guard = c->elem_guard || !(pp.K / c->rho_min > 0.0 && pp.K / c->rho_min <= DBL_MAX);
  return pp;
}

// J_local = vol/vol0 enters the pressure only through these coefficients or
// the terms of `full`
__host__ __device__ inline bool press_needs_jlocal(const PressParams &pp) {
  return pp.full || pp.alpha_free != 0.0 || pp.hg_coeff_free != 0.0 ||
         (pp.cf && (pp.alpha_contact != 0.0 || pp.hg_coeff_contact != 0.0));
}

// Contact detection of calcElemPressure (Mechanical.C:729-745): an element is
// in contact when one of its nodes carries a contact force of positive
// squared length (the last step's: contforce is never reset).
template <int NN, class NodeId>
__device__ __forceinline__ bool elem_in_contact(const PressParams &pp, NodeId node) {
  bool in = false;
  if (pp.cf) {
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const long nid = node(a);
      const double cx = pp.cf[nid], cy = pp.cf[(long)pp.N + nid], cz = pp.cf[2L * pp.N + nid];
      in = in || (cx * cx + cy * cy + cz * cz > 0);
    }
  }
  return in;
}

// An element whose volume is negative or not finite (inverted or broken):
// the reference still evaluates every stabilisation and viscosity term there,
// and h = vol^(1/3) is NaN, so zero coefficients give 0 * NaN = NaN.
__host__ __device__ inline bool vol_regular(double vol) { return vol >= 0.0 && vol <= DBL_MAX; }

// Elements of the pressure's general evaluation: every term computed as the
// reference writes it, whatever the coefficients.  Elements outside it have
// finite J_local, J_avg, h, c and tau, where a zero coefficient makes its
// term exactly zero.
__device__ __forceinline__ bool press_general(const PressParams &pp, double vol, double J_avg) {
  return pp.guard || !vol_regular(vol) || !(fabs(J_avg) <= DBL_MAX);
}

// The element pressure of calcElemPressure (Mechanical.C:746-815) from the
// element's averaged nodal ratio J_avg.  rho() and div_v() (the velocity
// divergence times detJ, sum_a dH_a . v_a) are evaluated only when the
// stabilisation terms need them or the element is in the general evaluation
// (`general`: the reference's 0 * NaN then makes the pressure NaN, as it
// does for an inverted element in compression).
template <class Rho, class DivV>
__device__ __forceinline__ double anp_elem_pressure(const PressParams &pp, bool is_contact, bool general,
                                                    double J_local, double J_avg, double vol, Rho rho,
                                                    DivV div_v_fn) {
  double al = is_contact ? pp.alpha_contact : pp.alpha_free;
  double J_bar = al * J_local + (1 - al) * J_avg;
  if (J_bar < pp.J_min) J_bar = 0.2;
  double lf = pp.log_factor;
  // log_factor 0: 0 * log(J_bar) is exactly 0 for a positive finite J_bar
  double lg = (lf != 0.0 || !(J_bar > 0.0 && J_bar <= DBL_MAX)) ? log(J_bar) : 0.0;
  double p_physical = -pp.K * (lf * lg + (1.0 - lf) * (J_bar - 1.0));
  double p_pspg = 0.0, p_q = 0.0;
  double p_hg = (is_contact ? pp.hg_coeff_contact : pp.hg_coeff_free) * pp.K * fabs(J_local - J_avg);
  if (pp.full || general) {
    double rho_e = rho();
    double h = pow(vol, 1.0 / 3.0);
    double c = sqrt(pp.K / rho_e);
    double tau = h / (2.0 * c);
    double div_v = div_v_fn();
    if (div_v < 0.0) {
      double x1 = pp.pspg_scale * tau * div_v * pp.K, x2 = pp.p_pspg_bulkfac * pp.K;
      p_pspg = (x2 < x1) ? x2 : x1;
      double q1 = pp.av_coeff_div * rho_e * h * c * (-div_v);
      double q2 = pp.av_coeff_bulk * pp.K * (1.0 - J_local);
      p_q = is_contact ? 0.5 * (q1 + q2) : ((q1 < q2) ? q2 : q1);
    }
  }
  return p_physical + p_pspg + p_hg + p_q;
}

// calcElemPressure element part (Mechanical.C:720-815).  With
// pspg/artificial-viscosity coefficients at zero the terms that need h, c and
// div v are exactly +-0 and are skipped (full == 0).
template <typename T, int NN>
__global__ __launch_bounds__(kBlock) void k_pressure_elem(int E, int N, const int *__restrict__ elnod,
                                                          const double *__restrict__ ratio,
                                                          const T *__restrict__ vol,
                                                          const T *__restrict__ vol0,
                                                          const T *__restrict__ rho,
                                                          const T *__restrict__ dH,
                                                          const double *__restrict__ v,
                                                          T *__restrict__ p, PressParams pp) {
  int e = gtid();
  if (e >= E) return;
  auto node = [&](int a) { return elnod[(long)a * E + e]; };
  const double ve = (double)vol[e], rho_e = (double)rho[e];
  double J_local = ve / (double)vol0[e];
  const bool is_contact = elem_in_contact<NN>(pp, node);
  double J_avg = 0.0;
#pragma unroll
  for (int a = 0; a < NN; ++a) J_avg += ratio[node(a)];
  J_avg /= NN;
  // this kernel reads J_local and rho anyway: the general evaluation is
  // decided per element (finite J_local, positive finite K/rho) instead of
  // by the scan the fused pass relies on
  const double kr = pp.K / rho_e;
  const bool general = !vol_regular(ve) || !(fabs(J_avg) <= DBL_MAX) || !(fabs(J_local) <= DBL_MAX) ||
                       !(kr > 0.0 && kr <= DBL_MAX);
  const double pe = anp_elem_pressure(
      pp, is_contact, general, J_local, J_avg, ve, [&] { return rho_e; },
      [&] {
        double div_v = 0.0;
#pragma unroll
        for (int a = 0; a < NN; ++a) {
          const int nid = node(a);
          div_v += (double)dH[(long)(0 * NN + a) * E + e] * v[nid] +
                   (double)dH[(long)(1 * NN + a) * E + e] * v[(long)N + nid] +
                   (double)dH[(long)(2 * NN + a) * E + e] * v[2L * N + nid];
        }
        return div_v;
      });
  p[e] = (T)pe;
}

// Temperature factors of the rate laws, evaluated once per element, or once
// per launch on the host when every element has the same temperature
struct RateTemp {
  double jc_soft;              // Johnson-Cook 1 - T_h^m
  double g_a, g_n, g_m, g_i;   // GMT C1 exp(C2 Tc), n1 Tc + n2, m1 Tc + m2, I1 Tc + I2
  double g_ni;                 // GMT -I1 Tc - I2
};

// the same expressions as the device's rate_temp_at (and the reference's)
__host__ __device__ inline double clamp_range(double v, double lo, double hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}
__host__ __device__ inline RateTemp rate_temp(double T_t, double T_m, double m_, double C1, double C2,
                                              double n1, double n2, double m1, double m2, double I1,
                                              double I2, double T_min, double T_max, double temp) {
  RateTemp r;
  const double T_h = (temp - T_t) / (T_m - T_t);
  r.jc_soft = 1.0 - pow(T_h, m_);
  const double Tc = clamp_range(temp, T_min, T_max);
  r.g_a = C1 * exp(C2 * Tc);
  r.g_n = n1 * Tc + n2;
  r.g_m = m1 * Tc + m2;
  r.g_i = I1 * Tc + I2;
  r.g_ni = -I1 * Tc - I2;
  return r;
}

struct MatParams {
  int model;
  double G, H, sy0, Kh, mh, eps0, eps1;
  double E;                                // Johnson-Cook tangent at zero plastic strain
  double B, n, C, eps_0, m, T_m, T_t;      // Johnson-Cook (A is the element's sy0)
  double C1, C2, n1, n2, m1, m2, I1, I2;   // GMT
  double e_min, e_max, er_min, er_max, T_min, T_max;
  double temp, max_edot;                   // uniform temperature, tangent strain-rate cap
  const double *Te;                        // per-element temperature, or null
  const double *Tn;                        // thermal coupling: nodal temperature, or null
  const int *elnod;                        //   with the connectivity [nn][Es]
  int nn, Es;
  int t_uniform;                           // every element at temp: ut holds the factors
  RateTemp ut;
};

// the rate laws' temperature factors at temperature temp
__device__ __forceinline__ RateTemp rate_temp_at(const MatParams &m, double temp) {
  return rate_temp(m.T_t, m.T_m, m.m, m.C1, m.C2, m.n1, m.n2, m.m1, m.m2, m.I1, m.I2, m.T_min, m.T_max, temp);
}

inline MatParams mat_params(const wfx_ctx *c) {
  const Material &m = c->prm.mat;
  const bool uniform = !c->temp_e && !c->thermal.on;
  return MatParams{m.model, m.G,     m.H,     m.sy0,   m.Kh,     m.mh,     m.eps0, m.eps1, m.E,
                   m.B,     m.n,     m.C,     m.eps_0, m.m,      m.T_m,    m.T_t,  m.C1,   m.C2,
                   m.n1,    m.n2,    m.m1,    m.m2,    m.I1,     m.I2,     m.e_min, m.e_max,
                   m.er_min, m.er_max, m.T_min, m.T_max, m.temp, m.max_edot, c->temp_e,
                   c->thermal.on ? c->nd.T : nullptr, c->ms.elnod, c->n, c->E, uniform ? 1 : 0,
                   rate_temp(m.T_t, m.T_m, m.m, m.C1, m.C2, m.n1, m.n2, m.m1, m.m2, m.I1, m.I2, m.T_min,
                             m.T_max, m.temp)};
}

// The temperature the hardening law of element e reads: with thermal
// coupling the mean of its nodal temperatures (the reference reads the nodal
// array at the element index, Mechanical.C:1725-1729, defined only for
// e < node count), else the element's or the uniform temperature.
__device__ __forceinline__ double law_temperature(const MatParams &m, int e) {
  if (m.Tn) {
    double t = 0.0;
    for (int a = 0; a < m.nn; ++a) t += m.Tn[m.elnod[(long)a * m.Es + e]];
    return t / m.nn;
  }
  return m.Te ? m.Te[e] : m.temp;
}

// The hardening laws evaluate in the element precision T: FP64 as the
// reference does, FP32-mixed in float (its element arithmetic is float).
__device__ __forceinline__ double rpow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float rpow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double rlog(double a) { return log(a); }
__device__ __forceinline__ float rlog(float a) { return logf(a); }
__device__ __forceinline__ double rexp(double a) { return exp(a); }
__device__ __forceinline__ float rexp(float a) { return expf(a); }
template <typename T>
__device__ __forceinline__ T rclamp(T v, double lo, double hi) {
  if (v < (T)lo) return (T)lo;
  if (v > (T)hi) return (T)hi;
  return v;
}

// effective strain rate of CalcStressStrain (Mechanical.C:1700-1705)
template <typename T>
__device__ __forceinline__ T eff_strain_rate(const T3<T> &D) {
  const T xx = D.xx, yy = D.yy, zz = D.zz, xy = D.xy, yz = D.yz, zx = D.xz;
  return sqrt(T(0.5) * ((xx - yy) * (xx - yy) + (yy - zz) * (yy - zz) + (zz - xx) * (zz - xx)) +
              T(3.0) * (xy * xy + yz * yz + zx * zx));
}

// std::min(edot, cap) (Mechanical.C:1737)
template <typename T>
__device__ __forceinline__ T cap_rate(const MatParams &m, T edot) {
  return ((T)m.max_edot < edot) ? (T)m.max_edot : edot;
}


// Kernels that update stresses are compiled twice: LAW_PLAIN handles the
// bilinear and Hollomon laws, LAW_RATE Johnson-Cook and GMT, so that the
// register budget of the common laws does not carry the rate laws' code.
enum { LAW_PLAIN = 0, LAW_RATE = 1 };
inline int law_class(int model) { return model >= JOHNSON_COOK ? LAW_RATE : LAW_PLAIN; }

// Values the tangent of a rate law shares with the yield stress of the same
// element and step (equal inputs give equal values, so they are reused)
template <typename T>
struct RateCache {
  T sr, lterm;  // Johnson-Cook: the rate and 1 + C ln(rate/eps_0)
  T er, pw_er;  // GMT: the clamped rate and er^(m1 Tc + m2)
  T ex_i;       // GMT: exp((I1 Tc + I2)/e)
};

// Calc*YieldStress (Material.cuh:353-364, :378-388, :412-441) at plastic
// strain ep and effective strain rate edot, the rate laws with the
// temperature factors rt; bilinear as sigma_y = sy0_e + H*ep (the reference
// leaves it undefined).  The clamps of GMT: Material.cuh:427-435, :451-459.
template <typename T, int LAW>
__device__ __forceinline__ T yield_stress(const MatParams &m, const RateTemp &rt, T sy0e, T ep, T edot,
                                          RateCache<T> &rc) {
  if constexpr (LAW == LAW_PLAIN) {
    if (m.model == HOLLOMON) {
      const T s = ep + (T)m.eps0;
      return (s > (T)m.eps1) ? (T)m.Kh * rpow(s, (T)m.mh) : (T)m.sy0;
    }
    return sy0e + (T)m.H * ep;
  } else if (m.model == JOHNSON_COOK) {
    rc.sr = (edot == T(0)) ? (T)1.e-5 : edot;
    rc.lterm = T(1) + (T)m.C * rlog(rc.sr / (T)m.eps_0);
    return (sy0e + (T)m.B * rpow(ep, (T)m.n)) * rc.lterm * (T)rt.jc_soft;
  } else {
    const T es = rclamp(ep, m.e_min, m.e_max);
    rc.er = rclamp(edot, m.er_min, m.er_max);
    rc.pw_er = rpow(rc.er, (T)rt.g_m);
    rc.ex_i = rexp((T)rt.g_i / es);
    return (T)rt.g_a * rpow(es, (T)rt.g_n) * rc.ex_i * rc.pw_er;
  }
}

// Calc*TangentModulus (Material.cuh:389-395, :397-411, :444-473); edot is
// the capped rate.
template <typename T, int LAW>
__device__ __forceinline__ T tangent_modulus(const MatParams &m, const RateTemp &rt, T ep, T edot,
                                             const RateCache<T> &rc) {
  if constexpr (LAW == LAW_PLAIN) {
    if (m.model == HOLLOMON) {
      const T s = ep + (T)m.eps0;
      return (s > (T)m.eps1) ? (T)(m.Kh * m.mh) * rpow(s, (T)(m.mh - 1.0)) : T(0);
    }
    return (T)m.H;
  } else if (m.model == JOHNSON_COOK) {
    if (ep > T(0)) {
      const T lterm = (edot == rc.sr) ? rc.lterm : T(1) + (T)m.C * rlog(edot / (T)m.eps_0);
      return (T)(m.n * m.B) * rpow(ep, (T)(m.n - 1.)) * lterm * (T)rt.jc_soft;
    }
    return (T)(m.E * 0.1);
  } else {
    const T es = rclamp(ep, m.e_min, m.e_max);
    const T er = rclamp(edot, m.er_min, m.er_max);
    const T pw_er = (er == rc.er) ? rc.pw_er : rpow(er, (T)rt.g_m);
    return (T)rt.g_a * pw_er * rpow(es, (T)(rt.g_n - 2.0)) * ((T)rt.g_ni + es * (T)rt.g_n) * rc.ex_i;
  }
}

// plastic heat source (Mechanical.C:1798-1810): frac * sigma : depdt with
// depdt = (1/dt) * strain_pl_incr
template <typename T>
__device__ __forceinline__ T plastic_heat(const T3<T> &sig, T i0, T i1, T i2, T i3, T i4, T i5, T dt,
                                          double frac) {
  const T r = T(1) / dt;
  return (T)frac * (sig.xx * (r * i0) + T(2) * sig.xy * (r * i3) + T(2) * sig.xz * (r * i5) +
                    sig.yy * (r * i1) + T(2) * sig.yz * (r * i4) + sig.zz * (r * i2));
}

// CalcStressStrain (Mechanical.C:1664-1839)
template <typename T, int LAW>
__global__ __launch_bounds__(kBlock) void k_stress_strain(
    int E, T dt, MatParams mp, const T *__restrict__ sr, const T *__restrict__ rr,
    const T *__restrict__ p, const T *__restrict__ sy0, const T *__restrict__ He, T *__restrict__ tau_io,
    T *__restrict__ eps_io, T *__restrict__ sigma, T *__restrict__ plincr,
    T *__restrict__ pl_strain, T *__restrict__ sigma_y, T *__restrict__ q_plheat, double plheat_frac) {
  int e = gtid();
  if (e >= E) return;
  const long Es = E;
  T3<T> tau = t3_from_sym(tau_io[0 * Es + e], tau_io[1 * Es + e], tau_io[2 * Es + e], tau_io[3 * Es + e],
                          tau_io[4 * Es + e], tau_io[5 * Es + e]);
  T3<T> D = t3_from_sym(sr[0 * Es + e], sr[1 * Es + e], sr[2 * Es + e], sr[3 * Es + e], sr[4 * Es + e],
                        sr[5 * Es + e]);
  T3<T> W = t3_from_antisym(rr[0 * Es + e], rr[1 * Es + e], rr[2 * Es + e]);
  T3<T> I = t3_identity<T>();
  T G = (T)mp.G;
  T3<T> Wt = t3_trans(W);
  T3<T> SRT = t3_mul(tau, Wt);
  T3<T> RS = t3_mul(W, tau);
  T3<T> cI = t3_scale(I, T(1.0 / 3.0) * t3_trace(D));
  T3<T> devD = t3_sub(D, cI);
  T3<T> A = t3_scale(devD, T(2) * G);
  T3<T> B = t3_add(A, SRT);
  T3<T> C = t3_add(B, RS);
  tau = t3_add(tau, t3_scale(C, dt));

  T pe = p[e];
  T3<T> pI = t3_scale(I, -pe);
  T3<T> st = t3_add(pI, tau);
  T3<T> s = t3_sub(st, t3_scale(I, T(1.0 / 3.0) * t3_trace(st)));
  T J2 = T(0.5) * (s.xx * s.xx + T(2) * s.xy * s.xy + T(2) * s.xz * s.xz + s.yy * s.yy +
                   T(2) * s.yz * s.yz + s.zz * s.zz);
  T sig_trial = sqrt(T(3) * J2);
  T ep = pl_strain[e];
  // bilinear plastic modulus: per element when given (specimen batches)
  const bool per_elem = LAW == LAW_PLAIN && mp.model == BILINEAR && He;
  const T edot = (LAW == LAW_RATE) ? eff_strain_rate(D) : T(0);
  RateTemp rt = mp.ut;
  if (LAW == LAW_RATE && !mp.t_uniform) rt = rate_temp_at(mp, law_temperature(mp, e));
  RateCache<T> rc;
  T sy = per_elem ? sy0[e] + He[e] * ep : yield_stress<T, LAW>(mp, rt, sy0[e], ep, edot, rc);
  sigma_y[e] = sy;
  T dep = T(0);
  if (sy < sig_trial) {
    T H = per_elem ? He[e] : tangent_modulus<T, LAW>(mp, rt, ep, cap_rate(mp, edot), rc);
    T dgamma = (sig_trial - sy) / (T(3) * G + H);
    T factor = T(1) - (T(3) * G * dgamma) / sig_trial;
    tau = t3_scale(s, factor);
    pl_strain[e] = ep + dgamma;
    dep = dgamma;
  }
  T3<T> sig = t3_add(pI, tau);
  if (plincr || q_plheat) {
    T i0 = 0, i1 = 0, i2 = 0, i3 = 0, i4 = 0, i5 = 0;
    if (dep > T(0)) {
      T f = dep / sy;
      i0 = f * (sig.xx - T(0.5) * (sig.yy + sig.zz));
      i1 = f * (sig.yy - T(0.5) * (sig.xx + sig.zz));
      i2 = f * (sig.zz - T(0.5) * (sig.xx + sig.yy));
      i3 = T(1.5) * f * (sig.xy);
      i4 = T(1.5) * f * (sig.yz);
      i5 = T(1.5) * f * (sig.xz);
    }
    if (plincr) {
      plincr[0 * Es + e] = i0; plincr[1 * Es + e] = i1; plincr[2 * Es + e] = i2;
      plincr[3 * Es + e] = i3; plincr[4 * Es + e] = i4; plincr[5 * Es + e] = i5;
    }
    if (q_plheat) q_plheat[e] = (dep > T(0)) ? plastic_heat<T>(sig, i0, i1, i2, i3, i4, i5, dt, plheat_frac) : T(0);
  }
  eps_io[0 * Es + e] = eps_io[0 * Es + e] + D.xx * dt;
  eps_io[1 * Es + e] = eps_io[1 * Es + e] + D.yy * dt;
  eps_io[2 * Es + e] = eps_io[2 * Es + e] + D.zz * dt;
  eps_io[3 * Es + e] = eps_io[3 * Es + e] + D.xy * dt;
  eps_io[4 * Es + e] = eps_io[4 * Es + e] + D.yz * dt;
  eps_io[5 * Es + e] = eps_io[5 * Es + e] + D.xz * dt;
  sigma[0 * Es + e] = sig.xx; sigma[1 * Es + e] = sig.yy; sigma[2 * Es + e] = sig.zz;
  sigma[3 * Es + e] = sig.xy; sigma[4 * Es + e] = sig.yz; sigma[5 * Es + e] = sig.xz;
  tau_io[0 * Es + e] = tau.xx; tau_io[1 * Es + e] = tau.yy; tau_io[2 * Es + e] = tau.zz;
  tau_io[3 * Es + e] = tau.xy; tau_io[4 * Es + e] = tau.yz; tau_io[5 * Es + e] = tau.xz;
}

// calcArtificialViscosity (Mechanical.C:1948-1977): the term added to each
// diagonal stress of an element with volumetric strain rate eps_v; false
// (nothing added) for abs(eps_v) <= 1e-12.  pow(y, 2) is y * y.
__device__ __forceinline__ bool artificial_viscosity_q(double K, double al, double be, double rho, double vol,
                                                       double eps_v, double &qs) {
  if (!(fabs(eps_v) > 1e-12)) return false;
  const double c = sqrt(K / rho);
  const double l = pow(vol, 1.0 / 3.0);
  const double y = eps_v * l;
  double q = al * c * fabs(eps_v) * l + be * (y * y);
  q = (1e9 < q) ? 1e9 : q;
  qs = (eps_v > 0) ? -q : q;
  return true;
}

// With both coefficients zero only the elements whose term can still be
// non-finite are visited (negative or non-finite volume, K/rho not positive
// and finite: the reference's 0 * NaN); the others add exactly zero.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_artificial_viscosity(int E, double K, double al,
                                                                 double be,
                                                                 const T *__restrict__ rho,
                                                                 const T *__restrict__ vol,
                                                                 const T *__restrict__ sr,
                                                                 T *__restrict__ sigma) {
  int e = gtid();
  if (e >= E) return;
  const long Es = E;
  const double ve = (double)vol[e], rho_e = (double)rho[e];
  if (al == 0.0 && be == 0.0) {
    const double kr = K / rho_e;
    if (vol_regular(ve) && kr > 0.0 && kr <= DBL_MAX) return;
  }
  const double eps_v = (double)sr[e] + (double)sr[Es + e] + (double)sr[2 * Es + e];
  double qs;
  if (artificial_viscosity_q(K, al, be, rho_e, ve, eps_v, qs))
    for (int k = 0; k < 3; ++k) sigma[k * Es + e] = (T)((double)sigma[k * Es + e] + qs);
}

// calcElemForces (Mechanical.C:375-481), 3D branch :448-457
template <typename T, int NN>
__global__ __launch_bounds__(kBlock) void k_elem_forces(int E, const T *__restrict__ dH,
                                                        const T *__restrict__ sigma,
                                                        T *__restrict__ f, T w) {
  int e = gtid();
  if (e >= E) return;
  const long Es = E;
  T s0 = sigma[e], s1 = sigma[Es + e], s2 = sigma[2 * Es + e], s3 = sigma[3 * Es + e],
    s4 = sigma[4 * Es + e], s5 = sigma[5 * Es + e];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    T h0 = dH[(long)(0 * NN + a) * Es + e], h1 = dH[(long)(1 * NN + a) * Es + e],
      h2 = dH[(long)(2 * NN + a) * Es + e];
    T f0 = T(0), f1 = T(0), f2 = T(0);
    f0 += h0 * s0 * T(1);
    f1 += h1 * s1 * T(1);
    f2 += h2 * s2 * T(1);
    f0 += h1 * s3 + h2 * s5;
    f1 += h0 * s3 + h2 * s4;
    f2 += h1 * s4 + h0 * s5;
    f[(long)(a * 3 + 0) * Es + e] = f0 * w;
    f[(long)(a * 3 + 1) * Es + e] = f1 * w;
    f[(long)(a * 3 + 2) * Es + e] = f2 * w;
  }
}

// 3D hexa viscous hourglass (f90_ver/src/Mechanical.f90:241-303): the
// reference C++ branch for 3D is empty (Mechanical.C:1845, :1939-1942).
__constant__ const signed char kHgGamma[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1},
                                                       {1, -1, -1, 1, -1, 1, 1, -1},
                                                       {1, -1, 1, -1, 1, -1, 1, -1},
                                                       {-1, 1, -1, 1, 1, -1, 1, -1}};
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hourglass(int E, int N, const int *__restrict__ elnod,
                                                      const double *__restrict__ v,
                                                      const T *__restrict__ vol,
                                                      const T *__restrict__ rho,
                                                      T *__restrict__ fhg, double coeff, double cs0) {
  int e = gtid();
  if (e >= E) return;
  const long Es = E;
  T V[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int g = elnod[(long)a * Es + e];
#pragma unroll
    for (int c = 0; c < 3; ++c) V[a][c] = (T)v[(long)c * N + g];
  }
  T hm[3][4];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) hm[c][j] = T(0);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) hm[c][j] = hm[c][j] + V[a][c] * (T)kHgGamma[j][a];
  T c_h = (T)(coeff * pow((double)vol[e], 0.6666666) * (double)rho[e] * 0.25 * cs0);
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T f = T(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) f = f - hm[c][j] * (T)kHgGamma[j][a];
      fhg[(long)(a * 3 + c) * Es + e] = f * c_h;
    }
}

// assemblyForces (Matrices.C:42-87): per node, CSR-ordered sum of the element
// forces, then CSR-ordered subtraction of the hourglass forces.  Non-finite
// results are set to zero and counted (Solver_explicit.C:777-785).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_assembly(int N, int E, const int *__restrict__ off,
                                                     const int *__restrict__ cnt,
                                                     const int *__restrict__ pad,
                                                     const int *__restrict__ pk,
                                                     const T *__restrict__ fel,
                                                     const T *__restrict__ fhg,
                                                     double *__restrict__ fi,
                                                     unsigned int *__restrict__ status) {
  int i = gtid();
  if (i >= N) return;
  const long Es = E;
  int o = off[i], c = cnt[i];
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  auto row3 = [&](const T *__restrict__ f) {
    return [=](int e, int a) {
      const long b = (long)a * 3;
      return double3{(double)f[(b + 0) * Es + e], (double)f[(b + 1) * Es + e], (double)f[(b + 2) * Es + e]};
    };
  };
  csr_walk(i, N, o, c, pad, pk, row3(fel), [&](double3 q) {
    f0 += q.x;
    f1 += q.y;
    f2 += q.z;
  });
  if (fhg)
    csr_walk(i, N, o, c, pad, pk, row3(fhg), [&](double3 q) {
      f0 -= q.x;
      f1 -= q.y;
      f2 -= q.z;
    });
  unsigned bad = 0;
  if (!isfinite(f0)) { f0 = 0.0; ++bad; }
  if (!isfinite(f1)) { f1 = 0.0; ++bad; }
  if (!isfinite(f2)) { f2 = 0.0; ++bad; }
  if (bad) atomicAdd(status, bad);
  fi[i] = f0;
  fi[(long)N + i] = f1;
  fi[2L * N + i] = f2;
}

// ---------------------------------------------------------------------------
// reductions (single work-group, fixed order: deterministic)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// reaction forces: one work-group per node group sums the fi component over
// the group's nodes (fixed order), out[g]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reaction(const int *__restrict__ goff,
                                                     const int *__restrict__ nodes,
                                                     const T *__restrict__ fcomp,
                                                     double *__restrict__ out) {
  __shared__ double sh[kBlock];
  const int g = blockIdx.x, k0 = goff[g], k1 = goff[g + 1];
  double s = 0.0;
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) s += (double)fcomp[nodes[k]];
  double r = block_sum(s, sh);
  if (threadIdx.x == 0) out[g] = r;
}

// calcMinEdgeLength (Domain_d.C:2224-2260): per-block minimum, then one block
__global__ __launch_bounds__(kBlock) void k_min_edge(int E, int N, int n, const int *__restrict__ elnod,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ partial) {
  __shared__ double sh[kBlock];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 1.0e6;
  if (e < E) {
    double P[4][3];
    for (int k = 0; k < 4; ++k) {
      int g = elnod[(long)k * E + e];
      for (int c = 0; c < 3; ++c) P[k][c] = x[(long)c * N + g];
    }
    const int ep[6][2] = {{1, 0}, {2, 0}, {3, 0}, {2, 1}, {3, 1}, {3, 2}};
    for (int i = 0; i < 6; ++i) {
      double ex = P[ep[i][0]][0] - P[ep[i][1]][0], ey = P[ep[i][0]][1] - P[ep[i][1]][1],
             ez = P[ep[i][0]][2] - P[ep[i][1]][2];
      double len = sqrt(ex * ex + ey * ey + ez * ez);
      if (len < m) m = len;
    }
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(kBlock) void k_min_final(int nb, const double *__restrict__ partial,
                                                      double *__restrict__ out) {
  __shared__ double sh[kBlock];
  double m = 1.0e6;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) m = fmin(m, partial[k]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// computeEnergies (Mechanical.C:2145-2185): per-block partial sums, then one
// block; out[0] = Ekin, out[1] = sum sigma:D*vol
template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy_partial(int N, int E, const double *__restrict__ v,
                                                           const double *__restrict__ m,
                                                           const T *__restrict__ sigma,
                                                           const T *__restrict__ sr,
                                                           const T *__restrict__ vol,
                                                           double *__restrict__ partial) {
  __shared__ double sh[kBlock];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ek = 0.0, ei = 0.0;
  if (i < N) {
    double vx = v[i], vy = v[(long)N + i], vz = v[2L * N + i];
    ek = 0.5 * m[i] * (vx * vx + vy * vy + vz * vz);
  }
  if (i < E) {
    const long Es = E;
    double sd = 0.0;
    for (int k = 0; k < 6; ++k) sd += (double)sigma[k * Es + i] * (double)sr[k * Es + i];
    ei = sd * (double)vol[i];
  }
  double a = block_sum(ek, sh);
  double b = block_sum(ei, sh);
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = a; partial[2 * blockIdx.x + 1] = b; }
}
__global__ __launch_bounds__(kBlock) void k_energy_final(int nb, const double *__restrict__ partial,
                                                         double *__restrict__ out) {
  __shared__ double sh[kBlock];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) { a += partial[2 * k]; b += partial[2 * k + 1]; }
  double ra = block_sum(a, sh);
  double rb = block_sum(b, sh);
  if (threadIdx.x == 0) { out[0] = ra; out[1] = rb; }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers (explicitly instantiated per precision)
// ---------------------------------------------------------------------------
template <typename T> DevElem<T> &elem(wfx_ctx *c);
template <> inline DevElem<double> &elem<double>(wfx_ctx *c) { return c->e64; }
template <> inline DevElem<float> &elem<float>(wfx_ctx *c) { return c->e32; }

template <typename T> hipError_t launch_elem_jacobians(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_elem_vol(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_elem_density(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_nodal_vol(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_nodal_mass(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_strain_rates(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_pressure(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_stress_strain(wfx_ctx *c, double dt, hipStream_t s);
template <typename T> hipError_t launch_artificial_viscosity(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_elem_forces(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_hourglass(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_assembly(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_energies(wfx_ctx *c, double *dev_out2, double *partial, hipStream_t s);

#define WFX_DEFINE_LAUNCHERS(T)                                                                    \
  template <> hipError_t launch_elem_jacobians<T>(wfx_ctx * c, hipStream_t s) {                    \
    auto &el = elem<T>(c);                                                                         \
    if (c->n == 8)                                                                                 \
      hipLaunchKernelGGL((k_elem_jacobians<T, 8>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,  \
                         c->N, c->ms.elnod, c->nd.x, el.detJ, el.dH, c->status);                   \
    else                                                                                           \
      hipLaunchKernelGGL((k_elem_jacobians<T, 4>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,  \
                         c->N, c->ms.elnod, c->nd.x, el.detJ, el.dH, c->status);                   \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_elem_vol<T>(wfx_ctx * c, hipStream_t s) {                          \
    auto &el = elem<T>(c);                                                                         \
    T w = (c->n == 4) ? T(1.0 / 6.0) : T(8.0);                                                     \
    hipLaunchKernelGGL((k_elem_vol<T>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, el.detJ,    \
                       el.vol, w);                                                                 \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_elem_density<T>(wfx_ctx * c, hipStream_t s) {                      \
    auto &el = elem<T>(c);                                                                         \
    hipLaunchKernelGGL((k_elem_density<T>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,         \
                       el.rho0, el.vol0, el.vol, el.rho);                                          \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_nodal_vol<T>(wfx_ctx * c, hipStream_t s) {                         \
    auto &el = elem<T>(c);                                                                         \
    hipLaunchKernelGGL((k_nodal_vol<T>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N,            \
                       c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel_pad, c->ms.nodel_pk,     \
                       el.vol, c->nd.voln, c->n);                                                  \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_nodal_mass<T>(wfx_ctx * c, hipStream_t s) {                        \
    auto &el = elem<T>(c);                                                                         \
    hipLaunchKernelGGL((k_nodal_mass<T>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N,           \
                       c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel_pad, c->ms.nodel_pk,     \
                       el.rho, c->nd.voln, c->nd.mdiag);                                           \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_strain_rates<T>(wfx_ctx * c, hipStream_t s) {                      \
    auto &el = elem<T>(c);                                                                         \
    if (c->n == 8)                                                                                 \
      hipLaunchKernelGGL((k_strain_rates<T, 8>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,    \
                         c->N, c->ms.elnod, c->nd.v, el.dH, el.detJ, el.str_rate, el.rot_rate);    \
    else                                                                                           \
      hipLaunchKernelGGL((k_strain_rates<T, 4>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,    \
                         c->N, c->ms.elnod, c->nd.v, el.dH, el.detJ, el.str_rate, el.rot_rate);    \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_pressure<T>(wfx_ctx * c, hipStream_t s) {                          \
    auto &el = elem<T>(c);                                                                         \
    hipLaunchKernelGGL((k_pressure_nodes<T>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N,       \
                       c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel_pad, c->ms.nodel_pk,     \
                       el.vol, el.vol0, c->nd.anp_ratio);                                          \
    const PressParams pp = press_params(c);                                                        \
    if (c->n == 8)                                                                                 \
      hipLaunchKernelGGL((k_pressure_elem<T, 8>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,   \
                         c->N, c->ms.elnod, c->nd.anp_ratio, el.vol, el.vol0,                      \
                         el.rho, el.dH, c->nd.v, el.p, pp);                                        \
    else                                                                                           \
      hipLaunchKernelGGL((k_pressure_elem<T, 4>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,   \
                         c->N, c->ms.elnod, c->nd.anp_ratio, el.vol, el.vol0,                      \
                         el.rho, el.dH, c->nd.v, el.p, pp);                                        \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_stress_strain<T>(wfx_ctx * c, double dt, hipStream_t s) {          \
    auto &el = elem<T>(c);                                                                         \
    const MatParams mp = mat_params(c);                                                            \
    auto kern = law_class(mp.model) == LAW_RATE ? k_stress_strain<T, LAW_RATE>                     \
                                                : k_stress_strain<T, LAW_PLAIN>;                   \
    hipLaunchKernelGGL(kern, dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, (T)dt, mp, el.str_rate, \
                       el.rot_rate, el.p, el.sy0, el.He, el.tau, el.eps, el.sigma,                 \
                       el.strain_pl_incr, el.pl_strain, el.sigma_y,                                \
                       c->thermal.on ? el.q_plheat : (T *)nullptr, c->thermal.plheat_frac);        \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_artificial_viscosity<T>(wfx_ctx * c, hipStream_t s) {              \
    const auto &st = c->prm.stab;                                                                  \
    auto &el = elem<T>(c);                                                                         \
    hipLaunchKernelGGL((k_artificial_viscosity<T>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, \
                       c->prm.mat.K, st.artifvisc0, st.artifvisc1, el.rho, el.vol, el.str_rate,    \
                       el.sigma);                                                                  \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_elem_forces<T>(wfx_ctx * c, hipStream_t s) {                       \
    auto &el = elem<T>(c);                                                                         \
    if (c->n == 8)                                                                                 \
      hipLaunchKernelGGL((k_elem_forces<T, 8>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,     \
                         el.dH, el.sigma, el.f_elem, T(8.0));                                      \
    else                                                                                           \
      hipLaunchKernelGGL((k_elem_forces<T, 4>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E,     \
                         el.dH, el.sigma, el.f_elem, T(1.0 / 6.0));                                \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_hourglass<T>(wfx_ctx * c, hipStream_t s) {                         \
    if (c->n != 8 || c->prm.stab.hg_coeff == 0.0) return hipSuccess;                               \
    auto &el = elem<T>(c);                                                                         \
    hipLaunchKernelGGL((k_hourglass<T>), dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, c->N,      \
                       c->ms.elnod, c->nd.v, el.vol, el.rho, el.f_hg, c->prm.stab.hg_coeff,        \
                       c->prm.mat.cs0);                                                            \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_assembly<T>(wfx_ctx * c, hipStream_t s) {                          \
    auto &el = elem<T>(c);                                                                         \
    const T *fhg = (c->n == 8 && c->prm.stab.hg_coeff != 0.0) ? el.f_hg : nullptr;                 \
    hipLaunchKernelGGL((k_assembly<T>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N, c->E,       \
                       c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel_pad, c->ms.nodel_pk,     \
                       el.f_elem, fhg, c->nd.fi, c->status);                                       \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  template <> hipError_t launch_energies<T>(wfx_ctx * c, double *out2, double *partial,            \
                                            hipStream_t s) {                                       \
    auto &el = elem<T>(c);                                                                         \
    int nb = nblocks(c->N > c->E ? c->N : c->E);                                                   \
    hipLaunchKernelGGL((k_energy_partial<T>), dim3(nb), dim3(kBlock), 0, s, c->N, c->E, c->nd.v,   \
                       c->nd.mdiag, el.sigma, el.str_rate, el.vol, partial);                       \
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, s, nb, partial, out2);            \
    return hipGetLastError();                                                                      \
  }

}  // namespace wfx
