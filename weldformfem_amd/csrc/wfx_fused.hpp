// Fused explicit step: the per-step Domain_d sequence regrouped into four
// kernels with one grid-wide dependency between each pair.
//
//   F1 element pass   x gather -> J -> vol                      (CalcElemVol)
//   F2 node pass      CSR sum of vol -> V_n, ANP ratio V_n/V0_n, nodal mass
//                                                  (CalcNodalVol,
//                                                   CalcNodalMassFromVol,
//                                                   calcElemPressure nodes)
//   F3 element pass   J and dH recomputed from x, strain rate and spin from
//                     the predicted v, ANP pressure, Jaumann + J2 return,
//                     internal and hourglass forces (f_elem - f_hg), added per
//                     node of the work-group's element cluster in LDS -> one
//                     partial sum per (cluster, node) slot
//                                                  (calcElemJAndDerivatives,
//                                                   calcElemStrainRates,
//                                                   calcElemPressure,
//                                                   CalcStressStrain,
//                                                   calcElemForces,
//                                                   calcElemHourglassForces)
//   F4 node pass      sum of the node's slots, acceleration, BCs,
//                     Chung-Hulbert correction, position update and the
//                     next step's prediction   (assemblyForces, calcAccel,
//                                                   ImposeBCA/BCV,
//                                                   UpdateCorrection*,
//                                                   UpdatePrediction)
//
// Every expression is the one of the per-step kernels, so the FP64 fused path
// differs from the per-step path only in the order of the nodal force sums:
// the hourglass force is subtracted per element instead of after the
// element-force sum, and a node's element forces are added per cluster
// (ascending element within a cluster, clusters in order) instead of in one
// ascending run.
//
// Device inputs beyond the per-step state:
//   nodel_pk[nnz]  packed CSR entry (element << 3) | local node
//   bc_slot[N]     -1 or index into bc_mask / bc_val (per-node Dirichlet data)
//   anp_ratio[N]   V_n / V0_n,   vsum0[N] = V0_n (constant after init)
#pragma once

#include "wfx_cluster.hpp"
#include "wfx_contact.hpp"
#include "wfx_kernels.hpp"

namespace wfx {
namespace {

// CSR entries fetched per batch of independent loads in the node passes
constexpr int kChunk = 8;  // also the width of the padded first-entries table
// Element-node rows of the per-step thermal rates (dTedt), tiled in blocks of
// 64 elements [E/64][n][64] so that a wavefront's rows are contiguous.
__device__ __forceinline__ long fidx(int row, long e, long Es, int nrow) {
  return (e >> 6) * ((long)nrow * 64) + (long)row * 64 + (e & 63);
}
// Element streams written once per step: non-temporal stores.
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}
// Element data the element pass reads once per step (connectivity, stress,
// strain, yield stress): non-temporal loads, so that they do not displace the
// nodal arrays its gathers reuse from L2 (measured: element pass -2%; the
// same loads in the volume and node passes made those slower, so they stay
// plain there).
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

// The nodal ANP ratio F2 leaves for the element pass: double, or (FP32-mixed)
// 1 + a float deviation
template <typename T>
__device__ __forceinline__ double anp_ratio_at(const double *ratio, int i) {
  if constexpr (std::is_same<T, double>::value) return ratio[i];
  else return 1.0 + (double)reinterpret_cast<const float *>(ratio)[i];
}

// V0_n: CSR-ordered sum of the initial volumes (calcElemPressure :698-705)
template <typename T>
__global__ __launch_bounds__(kBlock) void f0_vsum0(int N, const int *__restrict__ off,
                                                   const int *__restrict__ cnt,
                                                   const int *__restrict__ nodel_pk,
                                                   const T *__restrict__ vol0,
                                                   double *__restrict__ vsum0) {
  int i = gtid();
  if (i >= N) return;
  int o = off[i], c = cnt[i];
  double s0 = 0.0;
  for (int k = 0; k < c; ++k) s0 += (double)vol0[nodel_pk[o + k] >> 3];
  vsum0[i] = s0;
}

// F2: nodal volume sum in CSR order; ratio for the ANP pressure and nodal
// mass (the nodal volume field itself is written once per run, in phase 5).
template <typename T, int BS>
__global__ __launch_bounds__(BS) void f2_nodal(int N, int nodxelem, const int *__restrict__ off,
                                                   const int *__restrict__ cnt,
                                                   const int *__restrict__ nodel_pk,
                                                   const int *__restrict__ pad,
                                                   const T *__restrict__ vol,
                                                   const T *__restrict__ rho,
                                                   const double *__restrict__ vsum0,
                                                   double *__restrict__ ratio,
                                                   double *__restrict__ mdiag, double rho_u) {
  int i = gtid();
  if (i >= N) return;
  const int c = ld_stream(&cnt[i]), o = c > kChunk ? off[i] : 0;
  // The table, count and V0_n are read once per step: non-temporal loads, so
  // that the element volumes the gathers reuse stay in L2 (measured on the
  // 200x200x100 block: 89 -> 80 us FP64).  The ratio is stored plainly: the
  // element pass gathers it next.
  // CSR entries are taken eight at a time with independent loads, then
  // summed in CSR order (the order of the reference's serial loop); the
  // first eight come from the padded [8][N] table.
  // The densities of the first chunk are loaded together with the volumes and
  // kept in registers (every hexa node and most tet nodes have <= 8
  // elements); with a uniform density (rho_u > 0) no density is loaded.
  double s = 0.0;
  double r0[kChunk];
  for (int k = 0; k < c; k += kChunk) {
    int e[kChunk];
    double vv[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j)
      e[j] = (k == 0) ? (ld_stream(&pad[(long)j * N + i]) >> 3) : ((k + j < c) ? (nodel_pk[o + k + j] >> 3) : -1);
    // every load is issued (entries past the count read the chunk's first
    // element) and only the node's own entries are added: conditional loads
    // would each be waited for on their own
#pragma unroll
    for (int j = 0; j < kChunk; ++j) vv[j] = (double)vol[e[j] >= 0 ? e[j] : e[0]];
    if (k == 0 && rho_u <= 0.0) {
#pragma unroll
      for (int j = 0; j < kChunk; ++j) r0[j] = (double)rho[e[j] >= 0 ? e[j] : e[0]];
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j)
      if (k + j < c) s += vv[j];
  }
  // the ANP ratio V_n / V0_n: FP64 stores it; FP32-mixed stores its
  // deviation from 1 as a float in the same buffer (the element pass adds the
  // 1 back; the deviation keeps the float's relative precision)
  const double v0 = ld_stream(&vsum0[i]);
  if constexpr (std::is_same<T, double>::value) ratio[i] = s / v0;
  else reinterpret_cast<float *>(ratio)[i] = (float)(s / v0 - 1.0);
  double vn = s;
  vn /= nodxelem;
  double dc = (double)c, mass = 0.0;
  if (rho_u > 0.0) {  // every term is the same value: formed once, added c times
    const double term = 1.0 * rho_u * vn / dc;
    for (int k = 0; k < c; ++k) mass += term;
    st_stream(&mdiag[i], mass);
    return;
  }
  for (int k = 0; k < c; k += kChunk) {
    double rr[kChunk];
    if (k == 0) {
#pragma unroll
      for (int j = 0; j < kChunk; ++j) rr[j] = r0[j];
    } else {
#pragma unroll
      for (int j = 0; j < kChunk; ++j) rr[j] = (k + j < c) ? (double)rho[nodel_pk[o + k + j] >> 3] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j)
      if (k + j < c) mass += 1.0 * rr[j] * vn / dc;
  }
  mdiag[i] = mass;
}

// Nodes with many elements (tetrahedral meshes: about 23 per node) are
// summed by kF2Lanes lanes each: lane q takes the node's CSR entries q,
// q + kF2Lanes, ... in order, kF2Width at a time with every load issued, and
// the lanes' partial sums are combined by a butterfly, which leaves the same
// total in every lane.  V0_n goes through the same function, so the ANP ratio
// is exactly 1 wherever vol equals vol0.  val(e) gives entry e's term.
constexpr int kF2Lanes = 8, kF2Width = 4;
template <int L, typename F>
__device__ __forceinline__ double lane_csr_sum(int o, int c, int q, const int *__restrict__ nodel_pk, F val) {
  constexpr int W = kF2Width;
  double s = 0.0;
  for (int k0 = 0; k0 < c; k0 += W * L) {
    int e[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int k = k0 + q + j * L;
      e[j] = nodel_pk[o + (k < c ? k : c - 1)] >> 3;
    }
    double v[W];
#pragma unroll
    for (int j = 0; j < W; ++j) v[j] = val(e[j]);
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (k0 + q + j * L < c) s += v[j];
  }
#pragma unroll
  for (int m = 1; m < L; m <<= 1) s += __shfl_xor(s, m, L);
  return s;
}

// V0_n with kF2Lanes lanes per node (grid of N * kF2Lanes threads)
template <typename T>
__global__ __launch_bounds__(kBlock) void f0_vsum0_lanes(int N, const int *__restrict__ off,
                                                         const int *__restrict__ cnt,
                                                         const int *__restrict__ nodel_pk,
                                                         const T *__restrict__ vol0,
                                                         double *__restrict__ vsum0) {
  constexpr int L = kF2Lanes;
  const long g = (long)xcd_block() * kBlock + threadIdx.x;
  const long n = g / L;
  const int q = (int)(g % L), i = (int)(n < N ? n : N - 1);  // every lane reaches the butterfly
  const int o = off[i], c = cnt[i];
  const double s0 = lane_csr_sum<L>(o, c, q, nodel_pk, [&](int e) { return (double)vol0[e]; });
  if (q == 0 && n < N) vsum0[i] = s0;
}

// F2 with kF2Lanes lanes per node: the nodal volume sum, ANP ratio and mass
// of f2_nodal
template <typename T, int BS>
__global__ __launch_bounds__(BS) void f2_nodal_lanes(int N, int nodxelem, const int *__restrict__ off,
                                                     const int *__restrict__ cnt,
                                                     const int *__restrict__ nodel_pk,
                                                     const T *__restrict__ vol, const T *__restrict__ rho,
                                                     const double *__restrict__ vsum0,
                                                     double *__restrict__ ratio, double *__restrict__ mdiag,
                                                     double rho_u) {
  constexpr int L = kF2Lanes;
  const long g = (long)xcd_block() * BS + threadIdx.x;
  const long n = g / L;
  const int q = (int)(g % L), i = (int)(n < N ? n : N - 1);  // every lane reaches the butterflies
  const int o = off[i], c = cnt[i];
  const double v0 = vsum0[i];  // loaded with the offsets, not after the sum
  const double s = lane_csr_sum<L>(o, c, q, nodel_pk, [&](int e) { return (double)vol[e]; });
  double vn = s;
  vn /= nodxelem;
  const double dc = (double)c;
  double mass = 0.0;
  if (rho_u > 0.0) {  // every term is the same value: formed once, added c times
    const double term = 1.0 * rho_u * vn / dc;
    for (int k = 0; k < c; ++k) mass += term;
  } else {
    mass = lane_csr_sum<L>(o, c, q, nodel_pk, [&](int e) { return 1.0 * (double)rho[e] * vn / dc; });
  }
  if (q != 0 || n >= N) return;
  if constexpr (std::is_same<T, double>::value) ratio[i] = s / v0;
  else reinterpret_cast<float *>(ratio)[i] = (float)(s / v0 - 1.0);
  mdiag[i] = mass;
}

struct FusedParams {
  double dt;
  MatParams mp;
  PressParams pp;     // pp.full handled as in k_pressure_elem
  int need_jlocal;    // J_local = vol/vol0 enters the pressure
  double av_al, av_be;  // artificial viscosity (calcArtificialViscosity); both zero: none
  void *avq;          //   its diagonal stress term per element (T[E]), for the sigma field
  double hg_coeff, cs0;
  double rho_u;       // > 0: every element density equals it (no density loads)
  const double *dtp;  // non-null: the step's dt from the device (per-step CFL time step)
  double plheat_frac; // thermal coupling: fraction of plastic work into heat
  double k_T, cp_T;   //   conductivity and specific heat (rates when mp.Tn is set)
  // thermal expansion (with thermal coupling; texp null: none): exp_T, the
  // element means of the last rates (from F1), and where this step's rates
  // go: flat index flat_base[e] + a below N, divided by the node's mass
  double exp_T;
  const void *texp;
  void *dte_flat;
  const int *flat_base;
  const double *mdiag;
  // store the element outputs no later pass of the step reads (pressure,
  // yield stress, plastic heat, viscosity term): on a run's last step
  int write_out;
};

// dH(d,a) = adj(J)(d,:) . dN_a/dxi from the cofactor matrix C (adj = C^T),
// written exactly as calcElemJAndDerivatives does (Domain_d.C:1843-1850,
// :1866-1887).  a is a compile-time constant after unrolling.
template <typename T, int NN>
__device__ __forceinline__ T shape_dH(const T (&C)[3][3], int d, int a) {
  const T a0 = C[0][d], a1 = C[1][d], a2 = C[2][d];
  if constexpr (NN == 8) {
    const T q = T(0.125);
    switch (a) {
      case 0: return q * (-a0 - a1 - a2);
      case 1: return q * (a0 - a1 - a2);
      case 2: return q * (a0 + a1 - a2);
      case 3: return q * (-a0 + a1 - a2);
      case 4: return q * (-a0 - a1 + a2);
      case 5: return q * (a0 - a1 + a2);
      case 6: return q * (a0 + a1 + a2);
      default: return q * (-a0 + a1 + a2);
    }
  } else {
    switch (a) {
      case 0: return -a0 - a1 - a2;
      case 1: return a0;
      case 2: return a1;
      default: return a2;
    }
  }
}

// Walsh modes of a hexahedron's nodal values y[a] (local order of
// Domain_d.C:1313-1346): m[k] = sum_a y[a] prod_{i in k} s_i(a), the bits of
// k naming the natural directions (1 xi, 2 eta, 4 zeta) and s_i(a) = +-1 the
// sign of node a's natural coordinate (the table of Domain_d.C:1824-1826).
// m[1], m[2], m[4] are the sums calcElemJAndDerivatives weighs with 0.125
// (the rows of J for positions, of the velocity gradient in natural
// coordinates for velocities); m[6], m[5], m[3], m[7] are the hourglass
// amplitudes of Gamma rows 0..3 (Mechanical.f90:259-263: eta zeta, xi zeta,
// xi eta, xi eta zeta).  Three butterfly stages over the nodes in binary
// position order (local 0, 1, 3, 2, 4, 5, 7, 6); m[0] is not formed.
template <typename T>
__device__ __forceinline__ void hexa_modes(const T (&y)[8], T (&m)[8]) {
  const T a0 = y[1] + y[0], b0 = y[1] - y[0], a1 = y[2] + y[3], b1 = y[2] - y[3];
  const T a2 = y[5] + y[4], b2 = y[5] - y[4], a3 = y[6] + y[7], b3 = y[6] - y[7];
  const T c0 = a1 + a0, d0 = a1 - a0, c1 = b1 + b0, d1 = b1 - b0;
  const T c2 = a3 + a2, d2 = a3 - a2, c3 = b3 + b2, d3 = b3 - b2;
  m[0] = T(0);
  m[4] = c2 - c0;
  m[1] = c3 + c1;
  m[5] = c3 - c1;
  m[2] = d2 + d0;
  m[6] = d2 - d0;
  m[3] = d3 + d1;
  m[7] = d3 - d1;
}
// The inverse: nodal values z[a] = sum_k q[k] prod_{i in k} s_i(a) for
// mode coefficients q[1..7] (q[0] = 0), the same three stages in reverse.
template <typename T>
__device__ __forceinline__ void hexa_synthesis(const T (&q)[8], T (&z)[8]) {
  const T u0 = -q[4], u4 = q[4], u1 = q[1] - q[5], u5 = q[1] + q[5];
  const T u2 = q[2] - q[6], u6 = q[2] + q[6], u3 = q[3] - q[7], u7 = q[3] + q[7];
  const T v0 = u0 - u2, v2 = u0 + u2, v1 = u1 - u3, v3 = u1 + u3;
  const T v4 = u4 - u6, v6 = u4 + u6, v5 = u5 - u7, v7 = u5 + u7;
  z[0] = v0 - v1;
  z[1] = v0 + v1;
  z[3] = v2 - v3;
  z[2] = v2 + v3;
  z[4] = v4 - v5;
  z[5] = v4 + v5;
  z[7] = v6 - v7;
  z[6] = v6 + v7;
}

// Cofactors C (adj(J) = C^T, Matrix.h:694-727) and the Sarrus determinant
// (Matrix.h:610-615) of J.  FP64 without contraction: the same roundings as
// the per-step Jacobian kernel, so that an element's volume at rest equals
// its initial volume bit for bit (the ANP ratio is then exactly 1 there).
#define WFX_COFACTORS(J, C, detJ)                                                                         \
  C[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]);                                                        \
  C[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);                                                       \
  C[0][2] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]);                                                        \
  C[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);                                                       \
  C[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]);                                                        \
  C[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);                                                       \
  C[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]);                                                        \
  C[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);                                                       \
  C[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]);                                                        \
  detJ = J[0][0] * J[1][1] * J[2][2] - J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2] +          \
         J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] - J[0][2] * J[1][1] * J[2][0];
template <typename T>
__device__ __forceinline__ void cofactors(const T (&J)[3][3], T (&C)[3][3], T &detJ) {
  if constexpr (std::is_same<T, double>::value) {
#pragma clang fp contract(off)
    WFX_COFACTORS(J, C, detJ)
  } else {
    WFX_COFACTORS(J, C, detJ)
  }
}
#undef WFX_COFACTORS

// pow(vol, 0.6666666) of the hourglass coefficient (Mechanical.f90:299): the
// FP64 path in double, FP32-mixed with the hardware base-2 logarithm and
// exponential (the same NaN, zero and infinity cases)
template <typename T>
__device__ __forceinline__ T vol_pow_hg(T vol) {
  if constexpr (std::is_same<T, double>::value) return pow(vol, 0.6666666);
  else return exp2f(0.6666666f * log2f(vol));
}

// The cluster's slots staged in LDS by the element pass: positions (double),
// predicted velocities (T), ANP ratio as F2 stored it (double, or the float
// deviation from 1) and node ids, [field][slot_cap].
template <typename T, int NN>
struct SlotLds {
  static constexpr int cap = slot_cap(NN);
  static constexpr int bytes = cap * (3 * 8 + 3 * (int)sizeof(T) + (int)sizeof(T) + 4);
  double *x;
  T *v, *r;
  int *node;
  __device__ explicit SlotLds(unsigned char *base)
      : x(reinterpret_cast<double *>(base)),
        v(reinterpret_cast<T *>(base + 3 * 8 * cap)),
        r(reinterpret_cast<T *>(base + (3 * 8 + 3 * (int)sizeof(T)) * cap)),
        node(reinterpret_cast<int *>(base + (3 * 8 + 4 * (int)sizeof(T)) * cap)) {}
};

// Slots each thread of a cluster pass stages (slot_cap = kSlotsPerThread *
// kCluster): the staging rows past the cluster's slot count are written with
// a repeat of its last slot and never read, so that no load or store of the
// staging waits behind a branch
template <int NN>
constexpr int kSlotsPerThread = slot_cap(NN) / kCluster;

// What the first part of the element chain leaves for the second: geometry,
// strain rate and spin, hourglass amplitudes, the nodal ANP average and,
// with thermal coupling, the element's temperature terms.
template <typename T, int NN, bool TH>
struct ElemGeo {
  T C[3][3];
  T detJ, volv;
  T s00, s11, s22, s01, s12, s02, w01, w12, w02;
  T hm[3][4];
  double J_avg, div_v;
  bool is_contact;
  int g[TH ? NN : 1];
  double tsum, gT0, gT1, gT2;
};

// The Jacobian of an element from its staged positions, as the per-step
// Jacobian kernel forms it (calcElemJAndDerivatives, Domain_d.C:1822-1889):
// hexahedron rows 0.125 * (the signed sum of the nodes in the reference's
// order), tetrahedron rows x_{i+1} - x_0; FP32-mixed from FP64 differences to
// node 0 narrowed to float.  FP64 without contraction, so that the volume
// at rest equals the initial volume bit for bit.
template <typename T, int NN>
__device__ __forceinline__ void staged_jacobian(const double *lx, int cap, const int (&ls)[NN], T (&J)[3][3]) {
  T X[NN][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double *xc = lx + c * cap;
    if constexpr (std::is_same<T, double>::value) {
#pragma unroll
      for (int a = 0; a < NN; ++a) X[a][c] = xc[ls[a]];
    } else {
      const double x0 = xc[ls[0]];
      X[0][c] = T(0);
#pragma unroll
      for (int a = 1; a < NN; ++a) X[a][c] = T(xc[ls[a]] - x0);
    }
  }
  if constexpr (std::is_same<T, double>::value) {
#pragma clang fp contract(off)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if constexpr (NN == 8) {
        J[0][c] = T(0.125) * (-X[0][c] + X[1][c] + X[2][c] - X[3][c] - X[4][c] + X[5][c] + X[6][c] - X[7][c]);
        J[1][c] = T(0.125) * (-X[0][c] - X[1][c] + X[2][c] + X[3][c] - X[4][c] - X[5][c] + X[6][c] + X[7][c]);
        J[2][c] = T(0.125) * (-X[0][c] - X[1][c] - X[2][c] - X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
      } else {
        J[0][c] = X[1][c] - X[0][c];
        J[1][c] = X[2][c] - X[0][c];
        J[2][c] = X[3][c] - X[0][c];
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if constexpr (NN == 8) {
        J[0][c] = T(0.125) * (-X[0][c] + X[1][c] + X[2][c] - X[3][c] - X[4][c] + X[5][c] + X[6][c] - X[7][c]);
        J[1][c] = T(0.125) * (-X[0][c] - X[1][c] + X[2][c] + X[3][c] - X[4][c] - X[5][c] + X[6][c] + X[7][c]);
        J[2][c] = T(0.125) * (-X[0][c] - X[1][c] - X[2][c] - X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
      } else {
        J[0][c] = X[1][c] - X[0][c];
        J[1][c] = X[2][c] - X[0][c];
        J[2][c] = X[3][c] - X[0][c];
      }
    }
  }
}

// F3, first part: everything that reads the staged nodal data of element
// `ls` (its nodes as slot indices): Jacobian, cofactors and determinant
// (calcElemJAndDerivatives), the velocity gradient and from it the strain
// rate and spin (calcElemStrainRates), the hourglass amplitudes, the ANP
// average of the nodal ratios (calcElemPressure) and the contact test.
// sum_a dH(d,a) y_a = sum_i C[i][d] (0.125 m_i(y)) for a hexahedron
// (dH(d,a) = 0.125 sum_i s_i(a) C[i][d], Domain_d.C:1843-1850), and
// sum_i C[i][d] (y_{i+1} - y_0) for a tetrahedron (:1866-1887).
template <typename T, int NN, bool TH>
__device__ __forceinline__ void element_geometry(const FusedParams &fp, const SlotLds<T, NN> &L,
                                                 const int (&ls)[NN], ElemGeo<T, NN, TH> &G) {
  T J[3][3];
  staged_jacobian<T, NN>(L.x, L.cap, ls, J);
  cofactors(J, G.C, G.detJ);
  const T w = (NN == 4) ? T(1.0 / 6.0) : T(8.0);
  G.volv = G.detJ * w;
  // velocity modes: Mv[i][c] = 0.125 m_i(v_c) / the tetrahedron's differences
  T Mv[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T *vc = L.v + c * L.cap;
    if constexpr (NN == 8) {
      T y[8], m[8];
#pragma unroll
      for (int a = 0; a < 8; ++a) y[a] = vc[ls[a]];
      hexa_modes(y, m);
      Mv[0][c] = m[1];
      Mv[1][c] = m[2];
      Mv[2][c] = m[4];
      G.hm[c][0] = m[6];
      G.hm[c][1] = m[5];
      G.hm[c][2] = m[3];
      G.hm[c][3] = m[7];
    } else {
      const T v0 = vc[ls[0]];
#pragma unroll
      for (int i = 0; i < 3; ++i) Mv[i][c] = vc[ls[i + 1]] - v0;
#pragma unroll
      for (int j = 0; j < 4; ++j) G.hm[c][j] = T(0);
    }
  }
  // Gv(d,c) = sum_a dH(d,a) v_a,c up to the hexahedron's 0.125; the
  // reference's D and W (Mechanical.C:58-116): f = 1/detJ, off-diagonals
  // halved
  T Gv[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) Gv[d][c] = G.C[0][d] * Mv[0][c] + G.C[1][d] * Mv[1][c] + G.C[2][d] * Mv[2][c];
  const T f = T(1) / G.detJ;
  const T q = (NN == 8) ? T(0.125) * f : f, qh = T(0.5) * q;
  G.s00 = q * Gv[0][0];
  G.s11 = q * Gv[1][1];
  G.s22 = q * Gv[2][2];
  G.s01 = qh * (Gv[1][0] + Gv[0][1]);
  G.w01 = qh * (Gv[1][0] - Gv[0][1]);
  G.s12 = qh * (Gv[2][1] + Gv[1][2]);
  G.w12 = qh * (Gv[2][1] - Gv[1][2]);
  G.s02 = qh * (Gv[2][0] + Gv[0][2]);
  G.w02 = qh * (Gv[2][0] - Gv[0][2]);
  // div v times detJ for the pressure's stabilisation terms
  G.div_v = (NN == 8 ? 0.125 : 1.0) * ((double)Gv[0][0] + (double)Gv[1][1] + (double)Gv[2][2]);
  double J_avg = 0.0;
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    if constexpr (std::is_same<T, double>::value) J_avg += L.r[ls[a]];
    else J_avg += 1.0 + (double)L.r[ls[a]];
  }
  G.J_avg = J_avg / NN;
  G.is_contact = elem_in_contact<NN>(fp.pp, [&](int a) { return L.node[ls[a]]; });
  // thermal coupling (ThermalCalcs, Thermal.C:27-115): the conduction matrix
  // applied through the element's temperature gradient,
  // (Kt T)_i = sum_d dH(d,i) G_d with G_d = k_T w / detJ sum_j dH(d,j) T_j;
  // the nodal temperatures also give the rate laws' temperature (their mean)
  G.tsum = G.gT0 = G.gT1 = G.gT2 = 0.0;
  if constexpr (TH) {
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      G.g[a] = L.node[ls[a]];
      const double ta = fp.mp.Tn[G.g[a]];
      G.tsum += ta;
      G.gT0 += (double)shape_dH<T, NN>(G.C, 0, a) * ta;
      G.gT1 += (double)shape_dH<T, NN>(G.C, 1, a) * ta;
      G.gT2 += (double)shape_dH<T, NN>(G.C, 2, a) * ta;
    }
    const double kf = fp.k_T / (double)G.detJ * (NN == 4 ? 1.0 / 6.0 : 8.0);
    G.gT0 *= kf;
    G.gT1 *= kf;
    G.gT2 *= kf;
  }
}

// F3, second part: thermal expansion, hourglass coefficient, ANP pressure,
// Jaumann update and J2 return (CalcStressStrain), artificial viscosity,
// the element's internal and hourglass forces (calcElemForces,
// calcElemHourglassForces: f - f_hg, what the assembly adds to each node),
// handed to sink(a, node, fx, fy, fz, thermal term).  Hexahedral element
// passes of the rate laws in FP64 keep the hourglass amplitudes and the
// cofactors in LDS across the stress update (stash: this thread's column,
// stride BS): the laws' transcendental functions otherwise push the pass
// past 256 VGPRs.
// The element's state the second part reads (CalcStressStrain's stress,
// strain, plastic strain and yield stress; the per-element plastic modulus
// of specimen batches): loaded by the work-group before it stages its
// slots, so that these loads are in flight together with the staging
template <typename T>
struct ElemState {
  T tau[6], eps[6], ep, sy0, He;
};
template <typename T>
__device__ __forceinline__ void load_elem_state(int e, int E, const T *__restrict__ tau_io,
                                                const T *__restrict__ eps_io, const T *__restrict__ pl_strain,
                                                const T *__restrict__ sy0, const T *__restrict__ He,
                                                ElemState<T> &S) {
  const long Es = E;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    S.tau[k] = ld_stream(&tau_io[k * Es + e]);
    S.eps[k] = ld_stream(&eps_io[k * Es + e]);
  }
  S.ep = ld_stream(&pl_strain[e]);
  S.sy0 = ld_stream(&sy0[e]);
  S.He = He ? ld_stream(&He[e]) : T(0);
}

template <typename T, int NN, int LAW>
constexpr bool kStashLDS = NN == 8 && LAW == LAW_RATE && std::is_same<T, double>::value;
// hm[3][4], c_h, C[3][3] and with thermal coupling the gradient term G[3]
template <bool TH>
constexpr int kStashVals = TH ? 25 : 22;

template <typename T, int NN, int LAW, int BS, bool TH, class Sink>
__device__ __forceinline__ void element_material(
    int e, int E, const FusedParams &fp, ElemGeo<T, NN, TH> &G, const ElemState<T> &S,
    const T *__restrict__ vol0, const T *__restrict__ rho, const T *__restrict__ He, T *__restrict__ tau_io,
    T *__restrict__ eps_io, T *__restrict__ p_out, T *__restrict__ pl_strain, T *__restrict__ sigma_y,
    T *__restrict__ q_plheat, double *stash, Sink &&sink) {
  const long Es = E;
  T s00 = G.s00, s11 = G.s11, s22 = G.s22, s01 = G.s01, s12 = G.s12, s02 = G.s02;
  if constexpr (TH) {  // calcThermalExpansion (Thermal.C:164-178)
    if (fp.texp) {
      const double sh = fp.exp_T * (double)static_cast<const T *>(fp.texp)[e];
      s00 = (T)((double)s00 - sh * 1.0);
      s11 = (T)((double)s11 - sh * 1.0);
      s22 = (T)((double)s22 - sh * 1.0);
      s01 = (T)((double)s01 - sh * 0.0);
      s12 = (T)((double)s12 - sh * 0.0);
      s02 = (T)((double)s02 - sh * 0.0);
    }
  }
  const T volv = G.volv;
  T c_h = T(0);
  if constexpr (NN == 8) {
    if (fp.hg_coeff != 0.0) {
      const double rho_e = fp.rho_u > 0.0 ? fp.rho_u : (double)rho[e];
      c_h = (T)(fp.hg_coeff * rho_e * 0.25 * fp.cs0) * vol_pow_hg<T>(volv);
    }
  }
  if constexpr (kStashLDS<T, NN, LAW>) {
    double *st = stash;
    if constexpr (TH) {
      st[22 * BS] = G.gT0;
      st[23 * BS] = G.gT1;
      st[24 * BS] = G.gT2;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) st[(4 * c + j) * BS] = G.hm[c][j];
    st[12 * BS] = c_h;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) st[(13 + 3 * r + q) * BS] = G.C[r][q];
  }
  // ---- ANP pressure (calcElemPressure element part) ----------------------
  const PressParams &pp = fp.pp;
  const bool general = press_general(pp, (double)volv, G.J_avg);
  double J_local = 0.0;
  if (fp.need_jlocal || general) J_local = (double)volv / (double)vol0[e];
  const double div_v = G.div_v;
  const T pe = (T)anp_elem_pressure(
      pp, G.is_contact, general, J_local, G.J_avg, (double)volv, [&] { return (double)rho[e]; },
      [&] { return div_v; });
  if (fp.write_out) p_out[e] = pe;

  // ---- Jaumann update and J2 return (CalcStressStrain) -------------------
  T dt = (T)(fp.dtp ? *fp.dtp : fp.dt);
  const MatParams &mp = fp.mp;
  T3<T> tau = t3_from_sym(S.tau[0], S.tau[1], S.tau[2], S.tau[3], S.tau[4], S.tau[5]);
  T3<T> D = t3_from_sym(s00, s11, s22, s01, s12, s02);
  T3<T> W = t3_from_antisym(G.w01, G.w12, G.w02);
  T3<T> I = t3_identity<T>();
  T Gm = (T)mp.G;
  T3<T> Wt = t3_trans(W);
  T3<T> SRT = t3_mul(tau, Wt);
  T3<T> RS = t3_mul(W, tau);
  T3<T> cI = t3_scale(I, T(1.0 / 3.0) * t3_trace(D));
  T3<T> devD = t3_sub(D, cI);
  T3<T> A = t3_scale(devD, T(2) * Gm);
  T3<T> B = t3_add(A, SRT);
  T3<T> Cc = t3_add(B, RS);
  tau = t3_add(tau, t3_scale(Cc, dt));
  T3<T> pI = t3_scale(I, -pe);
  T3<T> st = t3_add(pI, tau);
  T3<T> s = t3_sub(st, t3_scale(I, T(1.0 / 3.0) * t3_trace(st)));
  T J2 = T(0.5) * (s.xx * s.xx + T(2) * s.xy * s.xy + T(2) * s.xz * s.xz + s.yy * s.yy +
                   T(2) * s.yz * s.yz + s.zz * s.zz);
  T sig_trial = sqrt(T(3) * J2);
  T ep = S.ep;
  // bilinear plastic modulus: per element when given (specimen batches)
  T sy, Hb = (T)mp.H;
  const bool per_elem = LAW == LAW_PLAIN && mp.model == BILINEAR && He;
  const T edot = (LAW == LAW_RATE) ? eff_strain_rate(D) : T(0);
  RateTemp rt = mp.ut;  // the rate laws' temperature factors (law_temperature on the loaded nodes)
  if constexpr (LAW == LAW_RATE) {
    if (!mp.t_uniform) {
      const double temp = TH ? G.tsum / NN : mp.Te[e];
      rt = rate_temp_at(mp, temp);
    }
  }
  RateCache<T> rc;
  if (per_elem) {
    Hb = S.He;
    sy = S.sy0 + Hb * ep;
  } else {
    sy = yield_stress<T, LAW>(mp, rt, S.sy0, ep, edot, rc);
  }
  if (fp.write_out) sigma_y[e] = sy;
  T dep = T(0);
  if (sy < sig_trial) {
    T H = per_elem ? Hb : tangent_modulus<T, LAW>(mp, rt, ep, cap_rate(mp, edot), rc);
    T dgamma = (sig_trial - sy) / (T(3) * Gm + H);
    T factor = T(1) - (T(3) * Gm * dgamma) / sig_trial;
    tau = t3_scale(s, factor);
    pl_strain[e] = ep + dgamma;
    dep = dgamma;
  }
  T3<T> sig = t3_add(pI, tau);
  double heat = 0.0;  // thermal coupling: the element's heat per node and cp
  if constexpr (TH) {  // plastic heat source of the step
    T q = T(0);
    if (dep > T(0)) {
      const T f = dep / sy;
      q = plastic_heat<T>(sig, f * (sig.xx - T(0.5) * (sig.yy + sig.zz)), f * (sig.yy - T(0.5) * (sig.xx + sig.zz)),
                          f * (sig.zz - T(0.5) * (sig.xx + sig.yy)), T(1.5) * f * (sig.xy), T(1.5) * f * (sig.yz),
                          T(1.5) * f * (sig.xz), dt, fp.plheat_frac);
    }
    if (fp.write_out) q_plheat[e] = q;
    heat = 0.9 * (double)q * (double)volv / NN / fp.cp_T;
  }
  // ---- calcArtificialViscosity (Mechanical.C:1948-1977), after the stress
  // update as in the reference's step (Solver_explicit.C:755-757): added to
  // the diagonal stresses the forces use; kept per element for the sigma field.
  // With zero coefficients it is zero except where the volume is negative or
  // not finite, or the scan flagged the densities (0 * NaN, as in the
  // reference); an infinite diagonal strain rate already makes the diagonal
  // stresses NaN.
  if (fp.av_al != 0.0 || fp.av_be != 0.0 || pp.guard || !vol_regular((double)volv)) {
    const double rho_e = fp.rho_u > 0.0 ? fp.rho_u : (double)rho[e];
    double qs = 0.0;
    if (artificial_viscosity_q(pp.K, fp.av_al, fp.av_be, rho_e, (double)volv,
                               (double)D.xx + (double)D.yy + (double)D.zz, qs)) {
      sig.xx = (T)((double)sig.xx + qs);
      sig.yy = (T)((double)sig.yy + qs);
      sig.zz = (T)((double)sig.zz + qs);
    }
    if (fp.write_out) static_cast<T *>(fp.avq)[e] = (T)qs;
  }
  st_stream(&eps_io[0 * Es + e], S.eps[0] + D.xx * dt);
  st_stream(&eps_io[1 * Es + e], S.eps[1] + D.yy * dt);
  st_stream(&eps_io[2 * Es + e], S.eps[2] + D.zz * dt);
  st_stream(&eps_io[3 * Es + e], S.eps[3] + D.xy * dt);
  st_stream(&eps_io[4 * Es + e], S.eps[4] + D.yz * dt);
  st_stream(&eps_io[5 * Es + e], S.eps[5] + D.xz * dt);
  st_stream(&tau_io[0 * Es + e], tau.xx); st_stream(&tau_io[1 * Es + e], tau.yy);
  st_stream(&tau_io[2 * Es + e], tau.zz); st_stream(&tau_io[3 * Es + e], tau.xy);
  st_stream(&tau_io[4 * Es + e], tau.yz); st_stream(&tau_io[5 * Es + e], tau.xz);

  T C[3][3], hm[3][4];
  double gT0 = G.gT0, gT1 = G.gT1, gT2 = G.gT2;
  if constexpr (kStashLDS<T, NN, LAW>) {  // back from LDS (volatile: real loads)
    volatile double *sv = stash;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) hm[c][j] = sv[(4 * c + j) * BS];
    c_h = sv[12 * BS];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) C[r][q] = sv[(13 + 3 * r + q) * BS];
    if constexpr (TH) {
      gT0 = sv[22 * BS];
      gT1 = sv[23 * BS];
      gT2 = sv[24 * BS];
    }
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) C[r][q] = G.C[r][q];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) hm[c][j] = G.hm[c][j];
  }
  // ---- internal forces (calcElemForces, Mechanical.C:448-457) and hourglass
  // f_c(a) = w sum_d dH(d,a) sigma(d,c): the cofactors times the stress give
  // the coefficients of the linear modes, Q(i,c) = 0.125 w sum_d C[i][d]
  // sigma(d,c) (0.125 w = 1 for the hexahedron); the hourglass force
  // (fi -= f_hg, Matrices.f90:641) adds c_h times the amplitudes to the
  // hourglass modes; one synthesis per component gives the nodal forces.
  const T sg0 = sig.xx, sg1 = sig.yy, sg2 = sig.zz, sg3 = sig.xy, sg4 = sig.yz, sg5 = sig.xz;
  T Q[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Q[i][0] = C[i][0] * sg0 + C[i][1] * sg3 + C[i][2] * sg5;
    Q[i][1] = C[i][0] * sg3 + C[i][1] * sg1 + C[i][2] * sg4;
    Q[i][2] = C[i][0] * sg5 + C[i][1] * sg4 + C[i][2] * sg2;
  }
  T fn[3][NN];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if constexpr (NN == 8) {
      T qm[8], z[8];
      qm[0] = T(0);
      qm[1] = Q[0][c];
      qm[2] = Q[1][c];
      qm[4] = Q[2][c];
      if (fp.hg_coeff != 0.0) {
        qm[6] = c_h * hm[c][0];
        qm[5] = c_h * hm[c][1];
        qm[3] = c_h * hm[c][2];
        qm[7] = c_h * hm[c][3];
      } else {
        qm[3] = qm[5] = qm[6] = qm[7] = T(0);
      }
      hexa_synthesis(qm, z);
#pragma unroll
      for (int a = 0; a < 8; ++a) fn[c][a] = z[a];
    } else {
      const T w = T(1.0 / 6.0);
      fn[c][1] = w * Q[0][c];
      fn[c][2] = w * Q[1][c];
      fn[c][3] = w * Q[2][c];
      fn[c][0] = -(fn[c][1] + fn[c][2] + fn[c][3]);
    }
  }
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    // the node's thermal term times its mass: -(Kt T)_a + heat / cp
    T th = T(0);
    int node = 0;
    if constexpr (TH) {
      const T h0 = shape_dH<T, NN>(C, 0, a), h1 = shape_dH<T, NN>(C, 1, a), h2 = shape_dH<T, NN>(C, 2, a);
      th = (T)(heat - ((double)h0 * gT0 + (double)h1 * gT1 + (double)h2 * gT2));
      node = G.g[a];
    }
    sink(a, node, fn[0][a], fn[1][a], fn[2][a], th);
  }
}

// Cluster tables of the fused element pass on the device (wfx_cluster.hpp)
struct ClusterArgs {
  const int *eoff;       // [C+1] first stored element of each cluster
  const int *sbase;      // [C+1] first slot of each cluster
  const uint16_t *loff;  // cluster c: its M+1 entry offsets at sbase[c] + c
  const uint16_t *ent;   // element-node entries t * 8 + a grouped by slot, cluster c at eoff[c] * n
  const int *snode;      // [S] slot -> node
  const uint16_t *lslot; // [E][n] element-node -> slot within its cluster
  const int *spos;       // [S] FP32-mixed: slot -> where its partial sum is stored (node-major)
  long S;                // slots: the row length of the partial sums
};

// Where the element pass stores a slot's partial sum.  FP64: in slot order
// (a cluster's sums are one contiguous run; the node pass reads a node's
// slots through their ids, and moves its bytes near the copy rate anyway).
// FP32-mixed: node-major, at the slot's place in its node's slot list (the
// node pass reads a node's sums contiguously from its offset, one dependent
// load less; measured on the 200x200x100 block: node pass 154 -> 118 us,
// element pass +12 us, while FP64 loses 15 us in the element pass).
template <typename T>
constexpr bool kNodeMajor = std::is_same<T, float>::value;

// Values per element-node the element pass adds up: the force and, with
// thermal coupling, the node's thermal term
template <bool TH>
constexpr int kNodeVals = TH ? 4 : 3;

// LDS of one work-group of the element pass, in turn: the cluster's staged
// slots (SlotLds), then the element-node values [kNodeVals * n][kCluster]
// and, in the same space (each thread only ever touches its own column, and
// reloads its stash before it writes its values), the per-thread stash of
// the FP64 hexahedral rate-law passes
template <typename T, int NN, int LAW, bool TH>
constexpr int kF3Vals = (kStashLDS<T, NN, LAW> && kStashVals<TH> * 8 > kNodeVals<TH> * NN * (int)sizeof(T))
                            ? kStashVals<TH> * 8 * kCluster
                            : kNodeVals<TH> * NN * (int)sizeof(T) * kCluster;
template <typename T, int NN, int LAW, bool TH>
constexpr int kF3Lds = kF3Vals<T, NN, LAW, TH> > SlotLds<T, NN>::bytes ? kF3Vals<T, NN, LAW, TH>
                                                                        : SlotLds<T, NN>::bytes;

// One element's nodes as slot indices of its cluster (a 16- or 8-byte load)
template <int NN>
__device__ __forceinline__ void load_lslot(const uint16_t *p, int (&ls)[NN]) {
  typedef unsigned int u32v __attribute__((ext_vector_type(NN / 2)));
  const u32v w = __builtin_nontemporal_load(reinterpret_cast<const u32v *>(p));
#pragma unroll
  for (int k = 0; k < NN / 2; ++k) {
    ls[2 * k] = (int)(w[k] & 0xffffu);
    ls[2 * k + 1] = (int)(w[k] >> 16);
  }
}

// F3: one work-group per element cluster.  The work-group first stages its
// cluster's slots (the distinct nodes of its elements: positions,
// velocities, ANP ratio, ids) in LDS with one load per slot and field; each
// thread then runs the element chain of one element from LDS and leaves its
// element-node values in the same LDS; then the work-group adds them per
// cluster node (the slot's entries, ascending element, in double) and writes
// one partial sum per slot (assemblyForces, Matrices.C:42-87, for the
// cluster's share of each node).
template <typename T, int NN, int LAW, bool TH>
__global__ __launch_bounds__(kCluster) void f3_cluster(
    int E, int N, FusedParams fp, ClusterArgs ca, const double *__restrict__ x, const T *__restrict__ v,
    const double *__restrict__ ratio, const T *__restrict__ vol0, const T *__restrict__ rho,
    const T *__restrict__ sy0, const T *__restrict__ He, T *__restrict__ tau_io, T *__restrict__ eps_io,
    T *__restrict__ p_out, T *__restrict__ pl_strain, T *__restrict__ sigma_y, T *__restrict__ part,
    T *__restrict__ q_plheat) {
  constexpr int R = kNodeVals<TH>;
  __shared__ __align__(16) unsigned char lds_raw[kF3Lds<T, NN, LAW, TH>];
  __shared__ uint32_t l_ent[kCluster * NN / 2];  // the cluster's entries, two per word
  T *lf = reinterpret_cast<T *>(lds_raw);
  const SlotLds<T, NN> L(lds_raw);
  const int c = xcd_block(), t = threadIdx.x;
  const int e0 = ca.eoff[c], ne = ca.eoff[c + 1] - e0;
  const int s0 = ca.sbase[c], M = ca.sbase[c + 1] - s0;  // M <= slot_cap(NN), checked at set-up
  const uint16_t *lo = ca.loff + s0 + c;
  // the nodes of this thread's slots t and t + kCluster: first, as the
  // staging loads depend on them
  constexpr int kSl = kSlotsPerThread<NN>;
  static_assert(kSl * kCluster == slot_cap(NN) && kSl <= 2, "one or two slots per thread");
  // Every load below is issued unconditionally, threads past the cluster's
  // elements, entries or slots repeating its last one (a cluster has at
  // least one element and NN slots), so that none is waited for on its own;
  // the values are used only where they are the thread's own.
  // First the nodes of this thread's slots t and t + kCluster, which the
  // staging loads depend on.
  int sn[kSl];
#pragma unroll
  for (int k = 0; k < kSl; ++k) sn[k] = ca.snode[s0 + min(t + k * kCluster, M - 1)];
  // The cluster's slot tables are loaded into registers before the element
  // chain, so that the loads complete under its arithmetic: the entries
  // (ne * NN uint16 from a 4-byte aligned start, since e0 * NN is even) go to
  // LDS after it; the entry offsets of the thread's own slots t and
  // t + kCluster stay in registers.
  const uint32_t *en32 = reinterpret_cast<const uint32_t *>(ca.ent + (long)e0 * NN);
  const int nw = ne * NN / 2;
  uint32_t ew[NN / 2];
#pragma unroll
  for (int k = 0; k < NN / 2; ++k) ew[k] = en32[min(t + k * kCluster, nw - 1)];
  uint32_t ow[kSl];  // (end << 16) | start of slots t and t + kCluster
  int sp[kSl];       //   and where their sums go
#pragma unroll
  for (int k = 0; k < kSl; ++k) {
    const int q = min(t + k * kCluster, M - 1);
    ow[k] = ((uint32_t)lo[q + 1] << 16) | lo[q];
    if constexpr (kNodeMajor<T>) sp[k] = ca.spos[s0 + q];
    else sp[k] = s0 + q;
  }
  int ls[NN];
  ElemState<T> S;
  {
    const int e = e0 + min(t, ne - 1);
    load_lslot<NN>(ca.lslot + (long)e * NN, ls);
    load_elem_state<T>(e, E, tau_io, eps_io, pl_strain, sy0, He, S);
  }
  // stage the slots: all loads first, then the LDS stores
  double sx[kSl][3];
  T sv[kSl][3], sr[kSl];
#pragma unroll
  for (int k = 0; k < kSl; ++k) {
    const int node = sn[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      sx[k][d] = x[(long)d * N + node];
      sv[k][d] = v[(long)d * N + node];
    }
    if constexpr (std::is_same<T, double>::value) sr[k] = ratio[node];
    else sr[k] = reinterpret_cast<const float *>(ratio)[node];
  }
#pragma unroll
  for (int k = 0; k < kSl; ++k) {
    const int j = t + k * kCluster;  // < slot_cap
    L.node[j] = sn[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      L.x[d * L.cap + j] = sx[k][d];
      L.v[d * L.cap + j] = sv[k][d];
    }
    L.r[j] = sr[k];
  }
  __syncthreads();
  ElemGeo<T, NN, TH> G;
  if (t < ne) element_geometry<T, NN, TH>(fp, L, ls, G);
  __syncthreads();  // the staged slots are read: their LDS takes the element-node values
  if (t < ne) {
    element_material<T, NN, LAW, kCluster, TH>(
        e0 + t, E, fp, G, S, vol0, rho, He, tau_io, eps_io, p_out, pl_strain, sigma_y, q_plheat,
        reinterpret_cast<double *>(lds_raw) + t, [&](int a, int node, T f0, T f1, T f2, T th) {
          lf[(a * R + 0) * kCluster + t] = f0;
          lf[(a * R + 1) * kCluster + t] = f1;
          lf[(a * R + 2) * kCluster + t] = f2;
          if constexpr (TH) {
            lf[(a * R + 3) * kCluster + t] = th;
            if (fp.dte_flat) {  // this step's rate at the flat index calcThermalExpansion reads
              const long q = (long)fp.flat_base[e0 + t] + a;
              if (q < N) static_cast<T *>(fp.dte_flat)[q] = (T)((double)th / fp.mdiag[node]);
            }
          }
        });
  }
#pragma unroll
  for (int k = 0; k < NN / 2; ++k) {
    const int w = t + k * kCluster;
    if (w < nw) l_ent[w] = ew[k];
  }
  __syncthreads();
  const uint16_t *en = reinterpret_cast<const uint16_t *>(l_ent);
#pragma unroll
  for (int r = 0; r < kSl; ++r) {  // M <= kSl * kCluster
    const int j = t + r * kCluster;
    if (j >= M) break;
    const int k0 = ow[r] & 0xffffu, k1 = ow[r] >> 16;
    double acc[R];
#pragma unroll
    for (int q = 0; q < R; ++q) acc[q] = 0.0;
    for (int k = k0; k < k1; ++k) {
      const int q = en[k];
      const int tt = q >> 3, a = q & 7;
#pragma unroll
      for (int u = 0; u < R; ++u) acc[u] += (double)lf[(a * R + u) * kCluster + tt];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if constexpr (kNodeMajor<T>) part[q * ca.S + sp[r]] = (T)acc[q];
      else st_stream(&part[q * ca.S + sp[r]], (T)acc[q]);
    }
  }
}

// F1 as a cluster pass: one work-group per element cluster stages its
// slots' positions (and node ids when thermal expansion reads them) in LDS,
// then each thread forms its element's Jacobian from LDS (the same Jacobian
// as the element pass's): vol = detJ * w (calcElemJAndDerivatives
// determinant + CalcElemVol), the inverted-element count, with minl_part the
// work-group's smallest edge (per-step CFL time step), with texp each
// element's mean of the last step's element-node temperature rates at the
// flat indices of its node ids (calcThermalExpansion's dTdt).
template <typename T, int NN>
__global__ __launch_bounds__(kCluster) void f1_cluster(int E, ClusterArgs ca, const double *__restrict__ x, int N,
                                                       T *__restrict__ vol, double *__restrict__ minl_part,
                                                       unsigned int *__restrict__ status,
                                                       const T *__restrict__ dte_flat, T *__restrict__ texp) {
  constexpr int cap = slot_cap(NN);
  __shared__ double lx[3 * cap];
  __shared__ int lnode[cap];
  const int c = xcd_block(), t = threadIdx.x;
  const int e0 = ca.eoff[c], ne = ca.eoff[c + 1] - e0;
  const int s0 = ca.sbase[c], M = ca.sbase[c + 1] - s0;  // M <= slot_cap(NN), checked at set-up
  // every load is issued unconditionally (threads past the cluster's
  // elements or slots repeat its last one), as in the element pass
  constexpr int kSl = kSlotsPerThread<NN>;
  int sn[kSl];
#pragma unroll
  for (int k = 0; k < kSl; ++k) sn[k] = ca.snode[s0 + min(t + k * kCluster, M - 1)];
  int ls[NN];
  load_lslot<NN>(ca.lslot + (long)(e0 + min(t, ne - 1)) * NN, ls);
  double sx[kSl][3];
#pragma unroll
  for (int k = 0; k < kSl; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) sx[k][d] = x[(long)d * N + sn[k]];
#pragma unroll
  for (int k = 0; k < kSl; ++k) {
    const int j = t + k * kCluster;  // < slot_cap
#pragma unroll
    for (int d = 0; d < 3; ++d) lx[d * cap + j] = sx[k][d];
    lnode[j] = sn[k];
  }
  __syncthreads();
  double medge = kMinEdgeInit;
  if (t < ne) {
    const int e = e0 + t;
    if (texp) {
      double m = 0.0;
#pragma unroll
      for (int a = 0; a < NN; ++a) m += 1.0 / NN * (double)dte_flat[lnode[ls[a]]];
      texp[e] = (T)m;
    }
    T J[3][3], C[3][3], det;
    staged_jacobian<T, NN>(lx, cap, ls, J);
    cofactors(J, C, det);
    const T w = (NN == 4) ? T(1.0 / 6.0) : T(8.0);
    vol[e] = det * w;  // CalcElemVol's 0 + detJ * w * 1 (the same value)
    count_inverted(det, status);
    if (minl_part) {  // the first four nodes, as calcMinEdgeLength takes them
      T P[4][3];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if constexpr (std::is_same<T, double>::value) P[a][d] = lx[d * cap + ls[a]];
          else P[a][d] = T(lx[d * cap + ls[a]] - lx[d * cap + ls[0]]);
        }
      medge = min_edge4<T>(P);
    }
  }
  if (minl_part) {
    medge = block_extreme<kCluster, true>(medge);
    if (threadIdx.x == 0) minl_part[blockIdx.x] = medge;
  }
}

// ---- thermal coupling ------------------------------------------------------
// ThermalCalcs (Thermal.C:27-115) without contact conduction.  Element e:
// Kt(i,j) = (sum_d dH(d,i) dH(d,j)) k_T / detJ * w, element-node rate
// -(Kt T)_i / m_i + 0.9 q_plheat vol / n / (m_i cp_T) into dTedt[i][E]; the
// node pass adds a node's rates in CSR order and advances T by dt.
template <typename T, int NN>
__device__ __forceinline__ void thermal_rates(int e, long Es, const int (&g)[NN], const double (&h)[3][NN],
                                              double detJ, double vol, double q, const double *__restrict__ Tn,
                                              const double *__restrict__ mdiag, double k_T, double cp_T,
                                              T *__restrict__ dTedt, long q0, int N, T *__restrict__ dte_flat) {
  const double w = (NN == 4) ? 1.0 / 6.0 : 8.0;
  double Te[NN];
#pragma unroll
  for (int a = 0; a < NN; ++a) Te[a] = Tn[g[a]];
  const double heat = 0.9 * q;
  const double elem_pow = heat * vol;
  const double pow_per_node = elem_pow / NN;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      double k = 0.0;
      k += h[0][i] * h[0][j];
      k += h[1][i] * h[1][j];
      k += h[2][i] * h[2][j];
      const double kt = k * k_T / detJ * w;
      acc += kt * Te[j];
    }
    const double m = mdiag[g[i]];
    const double m_inv = 1.0 / m;
    double r = -m_inv * acc;
    r += pow_per_node / (m * cp_T);
    dTedt[fidx(i, e, Es, NN)] = (T)r;
    if (dte_flat && q0 + i < N) dte_flat[q0 + i] = (T)r;  // the entries thermal expansion reads
  }
}

// per-step API: shape derivatives and determinant of this step's Jacobians
template <typename T, int NN>
__global__ __launch_bounds__(kBlock) void k_thermal_elem(
    int E, const int *__restrict__ elnod, const T *__restrict__ dH, const T *__restrict__ detJ,
    const T *__restrict__ vol, const T *__restrict__ q, const double *__restrict__ Tn,
    const double *__restrict__ mdiag, double k_T, double cp_T, T *__restrict__ dTedt,
    const int *__restrict__ flat_base, int N, T *__restrict__ dte_flat) {
  const int e = gtid();
  if (e >= E) return;
  int g[NN];
  double h[3][NN];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    g[a] = elnod[(long)a * E + e];
#pragma unroll
    for (int d = 0; d < 3; ++d) h[d][a] = (double)dH[(long)(d * NN + a) * E + e];
  }
  thermal_rates<T, NN>(e, E, g, h, (double)detJ[e], (double)vol[e], (double)q[e], Tn, mdiag, k_T, cp_T, dTedt,
                       dte_flat ? (long)flat_base[e] : 0L, N, dte_flat);
}

// calcThermalExpansion (Thermal.C:164-178): the strain rate loses (exp_T times
// the mean of the last step's element-node temperature rates at the flat
// indices given by the element's node ids) on its diagonal; the shear rates
// lose that times 0.
template <typename T, int NN>
__global__ __launch_bounds__(kBlock) void k_thermal_expansion(int E, const int *__restrict__ elnod,
                                                              const T *__restrict__ dte_flat, double exp_T,
                                                              T *__restrict__ sr) {
  const int e = gtid();
  if (e >= E) return;
  const long Es = E;
  double m = 0.0;
#pragma unroll
  for (int i = 0; i < NN; ++i) m += 1.0 / NN * (double)dte_flat[elnod[(long)i * Es + e]];
  const double sh = exp_T * m;
#pragma unroll
  for (int k = 0; k < 6; ++k) sr[k * Es + e] = (T)((double)sr[k * Es + e] - sh * (k < 3 ? 1.0 : 0.0));
}

// T += (sum of the node's element-node rates in CSR order) * dt; the rates
// sit in the element-force layout (fidx) and are gathered like the forces
template <typename T>
__global__ __launch_bounds__(kBlock) void k_thermal_node(int N, long Es, int nn, const int *__restrict__ off,
                                                         const int *__restrict__ cnt,
                                                         const int *__restrict__ nodel_pk,
                                                         const int *__restrict__ pad,
                                                         const T *__restrict__ dTedt, double dt,
                                                         double *__restrict__ Tn,
                                                         const double *__restrict__ qc,
                                                         const double *__restrict__ mdiag, double cp_T) {
  const int i = gtid();
  if (i >= N) return;
  const int c = cnt[i], o = c > kChunk ? off[i] : 0;
  double dTdt = 0.0;
  for (int k = 0; k < c; k += kChunk) {
    int pk[kChunk];
    double q[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j)
      pk[j] = (k == 0) ? pad[(long)j * N + i] : ((k + j < c) ? nodel_pk[o + k + j] : -1);
#pragma unroll
    for (int j = 0; j < kChunk; ++j)
      q[j] = (pk[j] >= 0) ? (double)dTedt[fidx(pk[j] & 7, pk[j] >> 3, Es, nn)] : 0.0;
#pragma unroll
    for (int j = 0; j < kChunk; ++j)
      if (k + j < c) dTdt += q[j];
  }
  // with the contact heat conduction of the last contact step (Thermal.C:124)
  const double qn = qc ? qc[i] * 1.0 / (mdiag[i] * cp_T) : 0.0;
  Tn[i] += (dTdt + qn) * dt;
}

struct NodeParams {
  double dt, alpha, beta, gamma;
  double f;        // 1/(1-alpha)
  double gdt;      // gamma*dt
  double bdt2;     // beta*dt*dt
  double pred_u;   // (0.5-beta)*dt   (next prediction)
  double pred_v;   // (1-gamma)*dt
  int predict;     // run the next step's prediction
  int write_fi;    // store the internal force (last step of a run, or its reaction is recorded)
  int read_ud;     // read the predicted increment (else recomputed from v' and prev_a)
  int write_ud;    // store the increment every step (else on a run's last step only)
  int run_first, run_last;  // FP64: first / last step of a run (both: u += u_dt directly)
  double *x_start;          //   positions when the run started
  float *x_lo;              //   rounding error of the position updates since then
  int has_fe;      // external forces present
  const double *dtp;  // per-step CFL time step: the step's dt on the device
  double *vmax_part;  //   and the per-block largest predicted speed
  const double *qcont;  // contact heat conduction per node (thermal with contact), or null
  double cp_T;
};

// the step-length terms of NodeParams for step length dt (same expressions
// as node_params on the host)
__device__ __forceinline__ void set_step_length(NodeParams &np, double dt) {
  np.dt = dt;
  np.gdt = np.gamma * dt;
  np.bdt2 = np.beta * dt * dt;
  np.pred_u = (0.5 - np.beta) * dt;
  np.pred_v = (1.0 - np.gamma) * dt;
}

// The state of node i the nodal update reads, loaded before the node's
// partial force sums so that these loads are in flight together with them.
template <typename T, bool CONTACT>
struct NodeIn {
  T v[3], pa[3], ud[3];
  float lo[3];
  double x[3], u[3], xs[3], fe[3], cf[3], m;
  int slot;
};
template <typename T, bool CONTACT>
__device__ __forceinline__ void load_node(int i, long Ns, const NodeParams &np, const double *__restrict__ cf,
                                          const double *__restrict__ fe, const double *__restrict__ mdiag,
                                          const int *__restrict__ bc_slot, const T *__restrict__ v,
                                          const T *__restrict__ prev_a, const T *__restrict__ u_dt,
                                          const double *__restrict__ x, const double *__restrict__ u,
                                          NodeIn<T, CONTACT> &s) {
  // row bases are uniform; the unsigned node index lets every row share one
  // 32-bit offset register
  const unsigned ui = (unsigned)i;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long r = d * Ns;
    s.v[d] = (v + r)[ui];
    s.pa[d] = (prev_a + r)[ui];
    // the predicted increment is used on a run's first step only; the other
    // steps load the array's first entry in every lane instead (one cache
    // line), so that no branch holds the load: a float widened inside one
    // waits for every load issued before it
    s.ud[d] = (u_dt + (np.read_ud ? r : 0L))[np.read_ud ? ui : 0u];
    s.x[d] = (x + r)[ui];
    s.u[d] = s.xs[d] = 0.0;
    s.lo[d] = 0.0f;
    if constexpr (std::is_same<T, double>::value) {  // FP32-mixed: u from x when a run ends
      // the error term is loaded every step (a run's first ignores it): a
      // load behind a branch would have its float widened inside the branch,
      // which waits for every load issued before it
      s.lo[d] = (np.x_lo + r)[ui];
      if (np.run_last) {
        s.u[d] = (u + r)[ui];
        s.xs[d] = (np.x_start + r)[ui];
      }
    }
    s.fe[d] = np.has_fe ? (fe + r)[ui] : 0.0;
    s.cf[d] = 0.0;
    if constexpr (CONTACT) s.cf[d] = (cf + r)[ui];
  }
  s.m = mdiag[ui];
  s.slot = bc_slot[ui];
}

// The nodal update of one step for node i given its assembled internal
// force f: non-finite scrub (Solver_explicit.C:777-785), calcAccel, ImposeBCA,
// UpdateCorrectionAccVel, ImposeBCV, UpdateCorrectionPos and the next step's
// UpdatePrediction + ImposeBCV.  FP64: every nodal array is double.
// FP32-mixed: velocity, previous acceleration, increment and internal force
// are stored in float (the update itself runs in double), positions stay
// double, and u is not touched per step (recovered as x - x_start when the
// run ends).  FP64 neither: positions advance as x += u_dt, and the exact
// rounding error of each of those additions (two-sum) is accumulated in a
// float beside them, so that u = u_start + ((x - x_start) + error) when the
// run ends is the sum of the increments to a few units in the last place of
// u (x - x_start alone loses the bits of x that u does not have); a one-step
// run adds the increment to u directly.
// The predicted increment u_dt = dt (v + (0.5 - beta) dt a) of the previous
// node pass is not stored between the steps of a run: it is formed again
// from what that pass stored, v' = v + (1 - gamma) dt a (or the imposed
// velocity, which the prediction then used itself) and prev_a = a, which
// gives it up to the rounding of v'.  The first step of a run reads the
// increment UpdatePrediction left (its v may differ from an imposed v'), a
// run's last step stores the corrected increment, and with a per-step time
// step (the previous step's length predicted it) it is read and stored
// every step.
template <typename T, bool CFL, bool CONTACT>
__device__ __forceinline__ double node_update(
    int i, long Ns, const NodeParams &np, double (&f)[3], const NodeIn<T, CONTACT> &s,
    const int *__restrict__ bc_mask, const double *__restrict__ bc_val, T *__restrict__ fi, T *__restrict__ v,
    T *__restrict__ prev_a, T *__restrict__ u_dt, double *__restrict__ x, double *__restrict__ u,
    unsigned int *__restrict__ status) {
  unsigned bad = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (!isfinite(f[d])) { f[d] = 0.0; ++bad; }
  if (bad) atomicAdd(status, bad);
  const int slot = s.slot;
  const int mask = slot >= 0 ? bc_mask[slot] : 0;
  const double m = s.m;
  double vs[3];
  const unsigned ui = (unsigned)i;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long r = d * Ns;
    T &fi_d = (fi + r)[ui], &v_d = (v + r)[ui], &pa_d = (prev_a + r)[ui], &ud_d = (u_dt + r)[ui];
    double &x_d = (x + r)[ui];
    if (np.write_fi) fi_d = (T)f[d];
    double a = (s.fe[d] - f[d]) / m;                    // calcAccel
    if constexpr (CONTACT) a += s.cf[d] / m;            //   + contact force
    if (mask & (1 << d)) a = 0.0;                       // ImposeBCA
    double pa = (double)s.pa[d];
    a = np.f * (a - np.alpha * pa);                     // UpdateCorrectionAccVel
    double vv = (double)s.v[d] + np.gdt * a;
    if (mask & (1 << d)) vv = bc_val[3L * slot + d];    // ImposeBCV
    // the previous node pass's prediction, from its v' and a
    double v0 = (double)s.v[d];
    if (!(mask & (1 << d))) v0 = v0 - np.pred_v * pa;
    double ud = np.read_ud ? (double)s.ud[d] : np.dt * (v0 + np.pred_u * pa);
    ud = ud + np.bdt2 * a;                              // UpdateCorrectionPos
    const double xo = s.x[d], xn = xo + ud;
    x_d = xn;
    if constexpr (std::is_same<T, double>::value) {
      const double bb = xn - xo;
      const double err = (xo - (xn - bb)) + (ud - bb);  // xo + ud - xn, exactly
      const double lo = (np.run_first ? 0.0 : (double)s.lo[d]) + err;
      if (np.run_last) {
        (u + r)[ui] = s.u[d] + (np.run_first ? ud : ((xn - s.xs[d]) + lo));
      } else {
        if (np.run_first) (np.x_start + r)[ui] = xo;
        (np.x_lo + r)[ui] = (float)lo;
      }
    }
    if (np.predict) {                                   // UpdatePrediction
      ud = np.dt * (vv + np.pred_u * a);
      vv = vv + np.pred_v * a;
      if (mask & (1 << d)) vv = bc_val[3L * slot + d];  // ImposeBCV
    }
    if (np.write_ud || !np.predict) ud_d = (T)ud;
    pa_d = (T)a;
    v_d = (T)vv;
    vs[d] = (double)(T)vv;
  }
  // speed of the predicted velocity, for the next step's CFL time step
  if constexpr (CFL) return sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
  return 0.0;
}

// Internal force of node i (and with TH its thermal term in f[3]): the sum of
// its slots' partial sums in ascending cluster order (assemblyForces,
// Matrices.C:42-87, completed across clusters), four at a time: FP64 through
// the slot ids of the node's list, FP32-mixed directly from its offset (the
// element pass stores those sums node-major).
template <typename T, bool TH = false>
__device__ __forceinline__ void gather_slots(int o, int c, const int *__restrict__ nsl, const T *__restrict__ part,
                                             long S, double (&f)[TH ? 4 : 3]) {
  constexpr int R = TH ? 4 : 3;
#pragma unroll
  for (int d = 0; d < R; ++d) f[d] = 0.0;
  // every load is issued (indices past the node's count repeat its last
  // slot) and only the node's own entries are added: conditional loads would
  // each be waited for on their own
  for (int k = 0; k < c; k += 4) {
    int sl[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = min(k + j, c - 1);
      if constexpr (kNodeMajor<T>) sl[j] = o + kk;
      else sl[j] = nsl[o + kk];
    }
    T q[4][R];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int d = 0; d < R; ++d) q[j][d] = part[d * S + sl[j]];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (k + j < c)
#pragma unroll
        for (int d = 0; d < R; ++d) f[d] += (double)q[j][d];
  }
}

// F4: assembly and the whole nodal update of one step; CFL: dt from the
// device and each block's largest predicted speed for the next step's dt;
// TH: thermal coupling, T += (sum of the node's thermal terms) / m * dt
template <typename T, int BS, bool CFL, bool CONTACT, bool TH>
__global__ __launch_bounds__(BS) void f4_node(
    int N, NodeParams np, const double *__restrict__ cf, const int *__restrict__ nsl_off,
    const int *__restrict__ nsl, const T *__restrict__ part, long S,
    const double *__restrict__ fe, const double *__restrict__ mdiag, const int *__restrict__ bc_slot,
    const int *__restrict__ bc_mask, const double *__restrict__ bc_val, T *__restrict__ fi,
    T *__restrict__ v, T *__restrict__ prev_a, T *__restrict__ u_dt,
    double *__restrict__ x, double *__restrict__ u, unsigned int *__restrict__ status,
    double *__restrict__ Tn) {
  int i = gtid();
  NodeParams q = np;
  if constexpr (CFL) set_step_length(q, *q.dtp);
  double speed = 0.0;
  if (i < N) {
    // the node's slot range first: the partial sums wait for it, not for
    // the state loads behind it
    const int o = nsl_off[i], c = nsl_off[i + 1] - o;
    NodeIn<T, CONTACT> st;
    load_node<T, CONTACT>(i, N, q, cf, fe, mdiag, bc_slot, v, prev_a, u_dt, x, u, st);
    double f[TH ? 4 : 3];
    gather_slots<T, TH>(o, c, nsl, part, S, f);
    speed = node_update<T, CFL, CONTACT>(i, N, q, *reinterpret_cast<double(*)[3]>(f), st, bc_mask, bc_val, fi, v,
                                         prev_a, u_dt, x, u, status);
    if constexpr (TH) {  // T += (dTdt + q_cont / (m cp)) dt (Thermal.C:124)
      double r = f[3] / st.m;
      if (q.qcont) r += q.qcont[i] * 1.0 / (st.m * q.cp_T);
      Tn[i] += r * q.dt;
    }
  }
  if constexpr (CFL) {
    if (q.predict) {
      speed = block_extreme<BS, false>(speed);
      if (threadIdx.x == 0) q.vmax_part[blockIdx.x] = speed;
    }
  }
}



// FP32-mixed fused runs keep v, prev_a, u_dt and fi in float between steps:
// narrowed once when the run starts, widened back (and u advanced by the
// position change) when it ends.
__global__ __launch_bounds__(kBlock) void f_pack(long n3, const double *__restrict__ v,
                                                 const double *__restrict__ pa,
                                                 const double *__restrict__ ud,
                                                 const double *__restrict__ x, float *__restrict__ v32,
                                                 float *__restrict__ pa32, float *__restrict__ ud32,
                                                 double *__restrict__ x0) {
  long i = gtid();
  if (i >= n3) return;
  v32[i] = (float)v[i];
  pa32[i] = (float)pa[i];
  ud32[i] = (float)ud[i];
  x0[i] = x[i];
}
__global__ __launch_bounds__(kBlock) void f_unpack(long n3, const float *__restrict__ v32,
                                                   const float *__restrict__ pa32,
                                                   const float *__restrict__ ud32,
                                                   const float *__restrict__ fi32,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ x0, double *__restrict__ v,
                                                   double *__restrict__ pa, double *__restrict__ a,
                                                   double *__restrict__ ud, double *__restrict__ fi,
                                                   double *__restrict__ u) {
  long i = gtid();
  if (i >= n3) return;
  v[i] = v32[i];
  pa[i] = a[i] = pa32[i];
  ud[i] = ud32[i];
  fi[i] = fi32[i];
  u[i] = u[i] + (x[i] - x0[i]);
}

// sigma = -p I + tau (CalcStressStrain :1766), for the field after a fused run
// (+ the step's artificial-viscosity term, calcArtificialViscosity, when avq)
// With zero viscosity coefficients (av_all false) the term is read only for
// the elements the element pass wrote it for: negative or non-finite volume
// in the last step (vol holds that step's volumes), or all under the guard.
template <typename T>
__global__ __launch_bounds__(kBlock) void f_sigma(int E, const T *__restrict__ p,
                                                  const T *__restrict__ tau, const T *__restrict__ avq,
                                                  const T *__restrict__ vol, int av_all,
                                                  T *__restrict__ sigma) {
  int e = gtid();
  if (e >= E) return;
  const long Es = E;
  T pe = p[e];
  T3<T> I = t3_identity<T>();
  T3<T> pI = t3_scale(I, -pe);
  T3<T> t = t3_from_sym(tau[0 * Es + e], tau[1 * Es + e], tau[2 * Es + e], tau[3 * Es + e], tau[4 * Es + e],
                        tau[5 * Es + e]);
  T3<T> s = t3_add(pI, t);
  if ((av_all || !vol_regular((double)vol[e])) && avq[e] != T(0)) {
    const double qs = (double)avq[e];
    s.xx = (T)((double)s.xx + qs);
    s.yy = (T)((double)s.yy + qs);
    s.zz = (T)((double)s.zz + qs);
  }
  sigma[0 * Es + e] = s.xx; sigma[1 * Es + e] = s.yy; sigma[2 * Es + e] = s.zz;
  sigma[3 * Es + e] = s.xy; sigma[4 * Es + e] = s.yz; sigma[5 * Es + e] = s.xz;
}

// Contact launchers (internal to each translation unit; the per-step API
// reaches them through the wrappers of wfx_k64.hip).
hipError_t launch_ext_face_areas(wfx_ctx *c, hipStream_t s) {
  const auto &ct = c->contact;
  hipLaunchKernelGGL(k_ext_face_areas, dim3(nblocks(ct.next)), dim3(kBlock), 0, s, ct.next, ct.ext_nodes,
                     ct.afc_off, ct.afc_face, ct.ext_face, c->N, c->nd.x, ct.node_area);
  return hipGetLastError();
}
hipError_t launch_elem_length(wfx_ctx *c, hipStream_t s) {
  hipLaunchKernelGGL(k_elem_length, dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, c->N, c->ms.elnod, c->nd.x,
                     c->contact.elem_length);
  return hipGetLastError();
}
// CalcContactForces with the predicted velocities v (T: the fused FP32 path's
// narrowed copy) and the host's dt or the device's (per-step CFL)
template <typename T>
hipError_t launch_contact(wfx_ctx *c, const T *v, bool device_dt, hipStream_t s) {
  const auto &ct = c->contact;
  ContactParams cp{ct.rt,        ct.rnode,  ct.rnormal, ct.rpplane, ct.rtri, ct.rmesh,
                   c->prm.mat.E, ct.pf,     ct.mu_sta,  ct.mu_dyn,  c->prm.dt,
                   device_dt ? c->scalars + kCflDt : nullptr,
                   ct.heat_cond, ct.T_const, c->nd.T, c->thermal.on ? ct.q_cont : nullptr};
  hipLaunchKernelGGL((k_contact<T>), dim3(nblocks(ct.next)), dim3(kBlock), 0, s, ct.next, ct.ext_nodes, c->N, cp,
                     c->nd.x, v, c->nd.mdiag, c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel, ct.elem_length,
                     ct.node_area, ct.contforce, ct.ut_prev, ct.mesh_in_contact, ct.cfn);
  return hipGetLastError();
}
hipError_t launch_rigid_update(wfx_ctx *c, bool device_dt, hipStream_t s) {
  auto &ct = c->contact;
  hipLaunchKernelGGL(k_rigid_update, dim3(1), dim3(kBlock), 0, s, ct.rn, ct.rt, ct.rnode, ct.rvel, ct.rtri,
                     ct.rnormal, ct.rpplane, c->prm.dt, device_dt ? c->scalars + kCflDt : nullptr);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_reaction(wfx_ctx *c, int dim, double *dev_out, hipStream_t s);
hipError_t launch_reaction_f32(wfx_ctx *c, const float *fi32, double *dev_out, hipStream_t s);

template <typename T>
hipError_t launch_vsum0(wfx_ctx *c, hipStream_t s) {
  if (c->f2_lanes) {
    hipLaunchKernelGGL((f0_vsum0_lanes<T>), dim3(nblocks((long)c->N * kF2Lanes)), dim3(kBlock), 0, s, c->N,
                       c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel_pk, elem<T>(c).vol0, c->nd.voln0_sum);
  } else {
    hipLaunchKernelGGL((f0_vsum0<T>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N, c->ms.nodel_offset,
                       c->ms.nodel_count, c->ms.nodel_pk, elem<T>(c).vol0, c->nd.voln0_sum);
  }
  return hipGetLastError();
}

inline NodeParams node_params(wfx_ctx *c, int predict) {
  const auto &P = c->prm;
  NodeParams np;
  np.dt = P.dt;
  np.alpha = P.alpha;
  np.beta = P.beta;
  np.gamma = P.gamma;
  np.f = 1.0 / (1.0 - P.alpha);
  np.gdt = P.gamma * P.dt;
  np.bdt2 = P.beta * P.dt * P.dt;
  np.pred_u = (0.5 - P.beta) * P.dt;
  np.pred_v = (1.0 - P.gamma) * P.dt;
  np.predict = predict;
  np.write_fi = c->fused_write_fi;
  np.read_ud = c->fused_read_ud;
  np.write_ud = c->fused_write_ud;
  np.run_first = c->fused_run_first;
  np.run_last = c->fused_run_last;
  np.x_start = c->x_start;
  np.x_lo = c->x_lo;
  np.has_fe = c->has_fe;
  np.dtp = nullptr;
  np.vmax_part = nullptr;
  np.qcont = (c->thermal.on && c->contact.on) ? c->contact.q_cont : nullptr;
  np.cp_T = c->thermal.cp_T;
  return np;
}

// dt of step `step` of a run from the partial extremes (per-step CFL time step)
template <typename T>
hipError_t launch_cfl_reduce(wfx_ctx *c, int nmin, int step, hipStream_t s) {
  hipLaunchKernelGGL((k_cfl_reduce<T>), dim3(1), dim3(kCflReduceBlock), 0, s, c->cfl_part, nmin,
                     c->cfl_part + c->cfl_cap_min, c->cfl_nvmax, elem<T>(c).rho, c->prm.mat.K, c->prm.cfl,
                     c->scalars, (step >= 0 && step < c->dt_cap) ? c->dt_hist : nullptr, step);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_thermal(wfx_ctx *c, hipStream_t s);

// Launch one phase (1..4) of a fused step on the current state with block
// size BS; launch_fused_phase picks BS from the context (autotuned or 256).
template <typename T, int NN, int LAW>
auto f3_pick(bool th) {
  return th ? f3_cluster<T, NN, LAW, true> : f3_cluster<T, NN, LAW, false>;
}

template <typename T, int BS, bool CFL, bool CT>
auto f4_pick(bool th) {
  return th ? f4_node<T, BS, CFL, CT, true> : f4_node<T, BS, CFL, CT, false>;
}

template <typename T, int BS>
hipError_t launch_fused_phase_bs(wfx_ctx *c, int phase, int predict, int step, hipStream_t s) {
  // per-step CFL time step inside a run; otherwise (step < 0) the host's dt
  const bool cfl = step >= 0 && c->prm.cfl > 0.0;
  const double *dtp = cfl ? c->scalars + kCflDt : nullptr;
  auto &el = elem<T>(c);
  const auto &P = c->prm;
  const auto &st = P.stab;
  const int N = c->N, E = c->E;
  const int gN = (N + BS - 1) / BS;
  // nodal arrays of the fused node pass in the element real type
  T *vN, *paN, *udN, *fiN;
  if constexpr (std::is_same<T, double>::value) {
    vN = c->nd.v; paN = c->nd.prev_a; udN = c->nd.u_dt; fiN = c->nd.fi;
  } else {
    vN = c->v32; paN = c->pa32; udN = c->ud32; fiN = c->fi32;
  }
  const ClusterArgs ca{c->ms.cl_eoff, c->ms.cl_sbase, c->ms.cl_loff,  c->ms.cl_ent,
                       c->ms.cl_snode, c->ms.cl_lslot, c->ms.spos, c->ms.nslot};
  if (phase == 1) {
    double *minl = cfl ? c->cfl_part : nullptr;
    const bool texp = c->thermal.on && c->thermal.exp_T != 0.0;
    auto kern = c->n == 8 ? f1_cluster<T, 8> : f1_cluster<T, 4>;
    hipLaunchKernelGGL(kern, dim3(c->ms.ncl), dim3(kCluster), 0, s, E, ca, c->nd.x, N, el.vol, minl, c->status,
                       el.dte_flat, texp ? el.texp : (T *)nullptr);
    if (cfl) {
      hipError_t e = launch_cfl_reduce<T>(c, c->ms.ncl, step, s);
      // calcMinEdgeLength also refreshes the element heights contact uses
      if (e == hipSuccess && c->contact.on) e = launch_elem_length(c, s);
      return e;
    }
  } else if (phase == 2) {
    if (c->f2_lanes) {
      const int gL = (int)(((long)N * kF2Lanes + BS - 1) / BS);
      hipLaunchKernelGGL((f2_nodal_lanes<T, BS>), dim3(gL), dim3(BS), 0, s, N, c->n, c->ms.nodel_offset,
                         c->ms.nodel_count, c->ms.nodel_pk, el.vol, el.rho, c->nd.voln0_sum, c->nd.anp_ratio,
                         c->nd.mdiag, c->rho_uniform);
    } else {
      hipLaunchKernelGGL((f2_nodal<T, BS>), dim3(gN), dim3(BS), 0, s, N, c->n, c->ms.nodel_offset,
                         c->ms.nodel_count, c->ms.nodel_pk, c->ms.nodel_pad, el.vol, el.rho, c->nd.voln0_sum,
                         c->nd.anp_ratio, c->nd.mdiag, c->rho_uniform);
    }
  } else if (phase == 3) {
    FusedParams fp;
    fp.dt = P.dt;
    fp.mp = mat_params(c);
    fp.pp = press_params(c);
    fp.need_jlocal = press_needs_jlocal(fp.pp);
    fp.av_al = st.artifvisc0;
    fp.av_be = st.artifvisc1;
    fp.avq = el.avq;
    fp.hg_coeff = (c->n == 8) ? st.hg_coeff : 0.0;
    fp.cs0 = P.mat.cs0;
    fp.rho_u = c->rho_uniform;
    fp.dtp = dtp;
    fp.plheat_frac = c->thermal.plheat_frac;
    fp.k_T = c->thermal.k_T;
    fp.cp_T = c->thermal.cp_T;
    const bool texp = c->thermal.on && c->thermal.exp_T != 0.0;
    fp.exp_T = c->thermal.exp_T;
    fp.texp = texp ? el.texp : nullptr;
    fp.dte_flat = texp ? el.dte_flat : nullptr;
    fp.flat_base = c->ms.flat_base;
    fp.mdiag = c->nd.mdiag;
    fp.write_out = c->fused_write_out;
    const bool rate = law_class(fp.mp.model) == LAW_RATE, th = c->thermal.on;
    auto kern = (c->n == 8) ? (rate ? f3_pick<T, 8, LAW_RATE>(th) : f3_pick<T, 8, LAW_PLAIN>(th))
                            : (rate ? f3_pick<T, 4, LAW_RATE>(th) : f3_pick<T, 4, LAW_PLAIN>(th));
    hipLaunchKernelGGL(kern, dim3(c->ms.ncl), dim3(kCluster), 0, s, E, N, fp, ca, c->nd.x, vN,
                       c->nd.anp_ratio, el.vol0, el.rho, el.sy0, el.He, el.tau, el.eps, el.p, el.pl_strain,
                       el.sigma_y, el.part, c->thermal.on ? el.q_plheat : (T *)nullptr);
    // contact forces after the element forces, before the assembly
    if (c->contact.on) return launch_contact<T>(c, vN, cfl, s);
  } else if (phase == 4) {
    NodeParams np = node_params(c, predict);
    np.dtp = dtp;
    np.vmax_part = cfl ? c->cfl_part + c->cfl_cap_min : nullptr;
    if (cfl && predict) c->cfl_nvmax = gN;
    const bool ct = c->contact.on, th = c->thermal.on;
    auto kern = cfl ? (ct ? f4_pick<T, BS, true, true>(th) : f4_pick<T, BS, true, false>(th))
                    : (ct ? f4_pick<T, BS, false, true>(th) : f4_pick<T, BS, false, false>(th));
    hipLaunchKernelGGL(kern, dim3(gN), dim3(BS), 0, s, N, np, c->contact.contforce, c->ms.nsl_off, c->ms.nsl,
                       el.part, c->ms.nslot, c->nd.fe, c->nd.mdiag, c->bc_slot, c->bc_mask, c->bc_tab, fiN, vN,
                       paN, udN, c->nd.x, c->nd.u, c->status, c->nd.T);
    // the rigid surfaces move after the position update
    if (c->contact.on && step >= 0) return launch_rigid_update(c, cfl, s);
  } else if (phase == 5) {
    const bool av = st.artifvisc0 != 0.0 || st.artifvisc1 != 0.0 || press_params(c).guard;
    hipLaunchKernelGGL((f_sigma<T>), dim3(nblocks(E)), dim3(kBlock), 0, s, E, el.p, el.tau, el.avq, el.vol,
                       (int)av, el.sigma);
    // the nodal volumes no step reads, once per run: the same CSR-order sum
    // and division as CalcNodalVol, so the field equals the per-step API's
    hipLaunchKernelGGL((k_nodal_vol<T>), dim3(nblocks(N)), dim3(kBlock), 0, s, N, c->ms.nodel_offset,
                       c->ms.nodel_count, c->ms.nodel_pad, c->ms.nodel_pk, el.vol, c->nd.voln, c->n);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// step >= 0: step index within a run (per-step CFL slots); -1: host dt
template <typename T>
hipError_t launch_fused_phase(wfx_ctx *c, int phase, int predict, hipStream_t s, int step) {
  const int bs = (phase >= 1 && phase <= 4) ? c->fused_bs[phase] : 256;
  switch (bs) {
    case 64: return launch_fused_phase_bs<T, 64>(c, phase, predict, step, s);
    case 128: return launch_fused_phase_bs<T, 128>(c, phase, predict, step, s);
    case 512: return launch_fused_phase_bs<T, 512>(c, phase, predict, step, s);
    default: return launch_fused_phase_bs<T, 256>(c, phase, predict, step, s);
  }
}

// Per-step CFL time step at the start of a run: the predicted velocities'
// per-block largest speeds for the first step's dt; packed: a fused run's
// velocities as its node passes store them (FP32-mixed: the float copy, as
// the later steps' speeds are taken)
template <typename T>
hipError_t launch_cfl_begin(wfx_ctx *c, hipStream_t s, bool packed) {
  if (packed && !std::is_same<T, double>::value)
    hipLaunchKernelGGL((k_cfl_max_speed<float>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N, c->v32,
                       c->cfl_part + c->cfl_cap_min);
  else
    hipLaunchKernelGGL((k_cfl_max_speed<double>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N, c->nd.v,
                       c->cfl_part + c->cfl_cap_min);
  c->cfl_nvmax = nblocks(c->N);
  return hipGetLastError();
}

// The time step of one per-step API step (after the prediction and the
// Jacobians, Solver_explicit.C:576-600) into the dt slot.
template <typename T>
hipError_t launch_cfl_step(wfx_ctx *c, int step, hipStream_t s) {
  hipError_t e = launch_cfl_begin<T>(c, s, false);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cfl_min_edge, dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, c->N, c->ms.elnod, c->nd.x,
                     c->cfl_part);
  return launch_cfl_reduce<T>(c, nblocks(c->E), -1, s);
}

// ThermalCalcs of one per-step API step: element rates from the stored
// Jacobians, then the nodal update with the host's dt.  (The fused path
// folds both into its element and node passes.)
template <typename T>
hipError_t launch_thermal(wfx_ctx *c, hipStream_t s) {
  auto &el = elem<T>(c);
  const auto &th = c->thermal;
  auto kern = c->n == 8 ? k_thermal_elem<T, 8> : k_thermal_elem<T, 4>;
  const bool texp = th.on && th.exp_T != 0.0;
  hipLaunchKernelGGL(kern, dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, c->ms.elnod, el.dH, el.detJ, el.vol,
                     el.q_plheat, c->nd.T, c->nd.mdiag, th.k_T, th.cp_T, el.dTedt, c->ms.flat_base, c->N,
                     texp ? el.dte_flat : (T *)nullptr);
  hipLaunchKernelGGL((k_thermal_node<T>), dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N, (long)c->E, c->n,
                     c->ms.nodel_offset, c->ms.nodel_count, c->ms.nodel_pk, c->ms.nodel_pad, el.dTedt, c->prm.dt,
                     c->nd.T, c->contact.on ? c->contact.q_cont : (const double *)nullptr, c->nd.mdiag, th.cp_T);
  return hipGetLastError();
}

// calcThermalExpansion of one per-step API step (no-op at exp_T = 0)
template <typename T>
hipError_t launch_thermal_expansion(wfx_ctx *c, hipStream_t s) {
  auto &el = elem<T>(c);
  if (c->thermal.exp_T == 0.0 || !el.dte_flat) return hipSuccess;
  auto kern = c->n == 8 ? k_thermal_expansion<T, 8> : k_thermal_expansion<T, 4>;
  hipLaunchKernelGGL(kern, dim3(nblocks(c->E)), dim3(kBlock), 0, s, c->E, c->ms.elnod, el.dte_flat, c->thermal.exp_T,
                     el.str_rate);
  return hipGetLastError();
}

// FP32-mixed runs keep the nodal velocity, previous acceleration and
// increment in float between the two; no-ops for FP64.
template <typename T>
hipError_t launch_fused_pack(wfx_ctx *c, hipStream_t s) {
  if constexpr (!std::is_same<T, double>::value) {
    const long n3 = 3L * c->N;
    hipLaunchKernelGGL(f_pack, dim3(nblocks(n3)), dim3(kBlock), 0, s, n3, c->nd.v, c->nd.prev_a, c->nd.u_dt,
                       c->nd.x, c->v32, c->pa32, c->ud32, c->x_start);
  }
  return hipGetLastError();
}
template <typename T>
hipError_t launch_fused_unpack(wfx_ctx *c, hipStream_t s) {
  if constexpr (!std::is_same<T, double>::value) {
    const long n3 = 3L * c->N;
    hipLaunchKernelGGL(f_unpack, dim3(nblocks(n3)), dim3(kBlock), 0, s, n3, c->v32, c->pa32, c->ud32, c->fi32,
                       c->nd.x, c->x_start, c->nd.v, c->nd.prev_a, c->nd.a, c->nd.u_dt, c->nd.fi, c->nd.u);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fused_steps(wfx_ctx *c, int nsteps, double *hist, hipStream_t s, hipEvent_t *ev) {
  constexpr bool mixed = !std::is_same<T, double>::value;
  {
    hipError_t e = launch_fused_pack<T>(c, s);
    if (e != hipSuccess) return e;
  }
  if (c->prm.cfl > 0.0) {
    hipError_t e = launch_cfl_begin<T>(c, s, true);
    if (e != hipSuccess) return e;
  }
  for (int k = 0; k < nsteps; ++k) {
    if (c->contact.on && (c->contact.step_count + k) % 10 == 0) {
      hipError_t e = launch_ext_face_areas(c, s);  // every 10 steps, at the step's start
      if (e != hipSuccess) return e;
    }
    // the internal force is observed after the run and by the reaction
    // history, the element outputs after the run
    c->fused_write_fi = (k + 1 == nsteps || hist) ? 1 : 0;
    c->fused_write_out = (k + 1 == nsteps) ? 1 : 0;
    // the increment is predicted again from v' and prev_a after the first
    // step, unless each step has its own length
    c->fused_read_ud = (k == 0 || c->prm.cfl > 0.0) ? 1 : 0;
    c->fused_write_ud = c->prm.cfl > 0.0 ? 1 : 0;
    c->fused_run_first = k == 0 ? 1 : 0;
    c->fused_run_last = k + 1 == nsteps ? 1 : 0;
    for (int ph = 1; ph <= 4; ++ph) {
      if (ev) {
        hipError_t e = hipEventRecord(ev[4 * k + ph - 1], s);
        if (e != hipSuccess) return e;
      }
      hipError_t e = launch_fused_phase<T>(c, ph, k + 1 < nsteps, s, k);
      if (e != hipSuccess) {
        c->fused_write_fi = c->fused_write_out = c->fused_read_ud = c->fused_write_ud = 1;
        c->fused_run_first = c->fused_run_last = 1;
        return e;
      }
    }
    c->fused_write_fi = c->fused_write_out = c->fused_read_ud = c->fused_write_ud = 1;
    c->fused_run_first = c->fused_run_last = 1;
    if (ev && k + 1 == nsteps) {
      hipError_t e = hipEventRecord(ev[4 * nsteps], s);
      if (e != hipSuccess) return e;
    }
    if (hist) {
      hipError_t e = mixed ? launch_reaction_f32(c, c->fi32, hist + (long)k * c->ngroups, s)
                           : launch_reaction(c, 2, hist + (long)k * c->ngroups, s);
      if (e != hipSuccess) return e;
    }
  }
  {
    hipError_t e = launch_fused_unpack<T>(c, s);
    if (e != hipSuccess) return e;
  }
  return launch_fused_phase<T>(c, 5, 0, s, -1);
}

}  // namespace wfx
