// Hand-written adjoint of the explicit step with checkpointing.
//
// The step map differentiated here is the fused step (wfx_step_fused) seen
// from one step boundary to the next.  The state at boundary k is
//   Z_k = (x_k, v'_k, ud_k, pa_{k-1}, tau_k, ep_k)
// with v'_k and ud_k the predicted velocity (boundary values imposed) and
// displacement increment of step k, pa_{k-1} the previous corrected
// acceleration.  Within step k (Solver_explicit.C:524-978):
//   element pass   vol, dH, D, W, p(r), Jaumann + radial return, forces   (wfx_adjoint.hpp)
//   node pass      r_i = S_i / S0_i, m_i = (S_i / n) * mean rho_i, S_i = sum vol_e
//                  fi = sum_e fnet, a = (fe - fi)/m, a = 0 where v is prescribed
//                  a'' = (a - alpha pa)/(1 - alpha), v'' = BCV(v' + gamma dt a'')
//                  ud' = ud + beta dt^2 a'', x_{k+1} = x_k + ud', pa_k = a''
//                  ud_{k+1} = dt (v'' + (0.5 - beta) dt a''),
//                  v'_{k+1} = BCV(v'' + (1 - gamma) dt a'')
// The loss cotangent enters through the reaction history R_{k,g} = sum over
// group g of fi_z after assembly.
//
// Forward: the fused step is run as usual and the state is stored every K
// steps (checkpoints).  Backward: for each segment, last first, the state is
// restored from its checkpoint, the K steps are recomputed while their
// inputs (x, v', tau, ep) and node-pass values (fi, m, r) are recorded, and
// the segment is then reversed step by step with five kernels:
//   a_node1   node-pass reverse: prediction, correction, BC masks, accel,
//             reaction cotangent -> fi and m cotangents
//   a_elem1   element reverse (recomputes the element forward first)
//   a_node2   gathers the r cotangents; volume cotangent per node
//   a_elem2   element volume cotangent -> positions
//   a_node3   gathers the element position/velocity cotangents
// All adjoint arithmetic is FP64; the forward runs in the context's precision.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/wfx.h"
#include "wfx_adjoint.hpp"
#include "wfx_internal.hpp"
#include "wfx_launch.hpp"

namespace wfx {

struct Adjoint {
  int nsteps = 0, every = 0, nck = 0, ng = 0, N = 0, E = 0, n = 0, precision = 0;
  bool recorded = false;
  size_t ck_stride = 0, slot_stride = 0;
  std::vector<void *> mem;
  double *ck = nullptr, *slot = nullptr, *hist = nullptr, *hbar = nullptr;
  double *xb = nullptr, *vb = nullptr, *udb = nullptr, *pab = nullptr;
  double *taub = nullptr, *epb = nullptr, *fib = nullptr, *mb = nullptr, *w = nullptr;
  double *contrib = nullptr, *volb = nullptr, *grad = nullptr;
  double *sy0d = nullptr, *Hd = nullptr, *rhod = nullptr, *rho_avg = nullptr, *sums = nullptr;
  int *rgroup = nullptr;
  std::vector<std::pair<void *, size_t>> final_bufs;
  std::vector<void *> final_saved;
  // per-step CFL time step: each step's dt (device and host), the node and
  // element shares of its cotangent, and the arg-min edge / arg-max speed
  bool cfl = false;
  double *dts = nullptr, *dtb = nullptr, *cfl_loc = nullptr;
  double *lawc = nullptr;  // rate laws at per-element temperatures: their constants per element [5][E]
  std::vector<double> dts_h;
};

void adjoint_free(wfx_ctx *c) {
  if (!c->adj) return;
  for (void *p : c->adj->mem) (void)hipFree(p);
  for (void *p : c->adj->final_saved) (void)hipFree(p);
  delete c->adj;
  c->adj = nullptr;
}

namespace {

constexpr int kB = 256;
inline int nb(long n) { return (int)((n + kB - 1) / kB); }
__device__ __forceinline__ long tid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }

// ---- state transfer --------------------------------------------------------
// checkpoint [x 3N][v 3N][ud 3N][pa 3N][tau 6E][ep E] (ud is read back by
// the run's first step only: the later ones form it from v and pa)
template <typename T>
__global__ void k_ck_save(long N, long E, const double *x, const T *v, const T *ud, const T *pa, const T *tau,
                          const T *ep, double *dst) {
  long i = tid();
  const long n3 = 3 * N;
  if (i < n3) dst[i] = x[i];
  else if (i < 2 * n3) dst[i] = (double)v[i - n3];
  else if (i < 3 * n3) dst[i] = (double)ud[i - 2 * n3];
  else if (i < 4 * n3) dst[i] = (double)pa[i - 3 * n3];
  else if (i < 4 * n3 + 6 * E) dst[i] = (double)tau[i - 4 * n3];
  else if (i < 4 * n3 + 7 * E) dst[i] = (double)ep[i - 4 * n3 - 6 * E];
}
template <typename T>
__global__ void k_ck_load(long N, long E, const double *src, double *x, T *v, T *ud, T *pa, T *tau, T *ep) {
  long i = tid();
  const long n3 = 3 * N;
  if (i < n3) x[i] = src[i];
  else if (i < 2 * n3) v[i - n3] = (T)src[i];
  else if (i < 3 * n3) ud[i - 2 * n3] = (T)src[i];
  else if (i < 4 * n3) pa[i - 3 * n3] = (T)src[i];
  else if (i < 4 * n3 + 6 * E) tau[i - 4 * n3] = (T)src[i];
  else if (i < 4 * n3 + 7 * E) ep[i - 4 * n3 - 6 * E] = (T)src[i];
}
// step record [x 3N][v 3N][fi 3N][m N][r N][tau 6E][ep E][pa 3N]; the first
// part before the step's kernels, the node-pass values after them
template <typename T>
__global__ void k_rec_pre(long N, long E, const double *x, const T *v, const T *tau, const T *ep, const T *pa,
                          double *dst) {
  long i = tid();
  const long n3 = 3 * N, o = 11 * N;
  if (i < n3) dst[i] = x[i];
  else if (i < 2 * n3) dst[i] = (double)v[i - n3];
  else if (i < 2 * n3 + 6 * E) dst[o + i - 2 * n3] = (double)tau[i - 2 * n3];
  else if (i < 2 * n3 + 7 * E) dst[o + i - 2 * n3] = (double)ep[i - 2 * n3 - 6 * E];
  else if (i < 3 * n3 + 7 * E) dst[o + i - 2 * n3] = (double)pa[i - 2 * n3 - 7 * E];
}
// (the ANP ratio as the fused node pass leaves it: double, or for FP32-mixed
// 1 + a float deviation in the same buffer)
template <typename T>
__global__ void k_rec_post(long N, const T *fi, const double *m, const double *r, double *dst) {
  long i = tid();
  const long n3 = 3 * N;
  if (i < n3) dst[6 * N + i] = (double)fi[i];
  else if (i < n3 + N) dst[9 * N + i - n3] = m[i - n3];
  else if (i < n3 + 2 * N) {
    const long k = i - n3 - N;
    if constexpr (std::is_same<T, double>::value) dst[10 * N + k] = r[k];
    else dst[10 * N + k] = 1.0 + (double)reinterpret_cast<const float *>(r)[k];
  }
}

__global__ void k_zero(long n, double *p) {
  long i = tid();
  if (i < n) p[i] = 0.0;
}

// element parameters in FP64 and per-node mean density
template <typename T>
__global__ void k_prep_elem(int E, const T *sy0, const T *He, double H, const T *rho, double *sy0d,
                            double *Hd, double *rhod) {
  int e = (int)tid();
  if (e >= E) return;
  sy0d[e] = (double)sy0[e];
  Hd[e] = He ? (double)He[e] : H;
  rhod[e] = (double)rho[e];
}
__global__ void k_prep_node(int N, const int *off, const int *cnt, const int *pk, const double *rhod,
                            double *rho_avg) {
  int i = (int)tid();
  if (i >= N) return;
  double s = 0.0;
  for (int k = 0; k < cnt[i]; ++k) s += rhod[pk[off[i] + k] >> 3];
  rho_avg[i] = cnt[i] > 0 ? s / cnt[i] : 0.0;
}

struct NodeAdj {
  double dt, alpha, beta, gamma;
  int has_fe, ng;
};

// node-pass reverse of step k; Z-bar of boundary k+1 in, of boundary k out
// (partial: the element gathers are added by a_node3).  dtb (per-step CFL
// time step, else null): node i's share of the step length's cotangent, from
// vv = v' + gamma dt ac, ud = ud + beta dt^2 ac and the prediction
// ud' = dt (vv + (0.5 - beta) dt ac), v'' = vv + (1 - gamma) dt ac, with ac the
// corrected acceleration (1/(1-alpha)) (a - alpha pa)
__global__ void a_node1(int N, int E, NodeAdj P, const double *rec, const double *fe, const int *bc_slot,
                        const int *bc_mask, const int *rgroup, const double *hbar_k, double *xb, double *vb,
                        double *udb, double *pab, double *fib, double *mb, double *dtb) {
  int i = (int)tid();
  if (i >= N) return;
  const long Ns = N;
  const double *fi = rec + 6 * Ns, m = rec[9 * Ns + i];
  const double *vrec = rec + 3 * Ns, *parec = rec + 11 * Ns + 7L * E;
  const int slot = bc_slot[i];
  const int mask = slot >= 0 ? bc_mask[slot] : 0;
  const double dt = P.dt, ia = 1.0 / (1.0 - P.alpha);
  double mbar = 0.0, dtbar = 0.0;
  for (int d = 0; d < 3; ++d) {
    const long id = d * Ns + i;
    const bool free_d = !((mask >> d) & 1);
    const double vp = vb[id], udp = udb[id], pa_in = pab[id], xbar = xb[id];
    const double vpm = free_d ? vp : 0.0;
    const double vpp = dt * udp + vpm;       // v'' after its boundary values
    const double vppr = free_d ? vpp : 0.0;  // v'' before them
    const double app = dt * dt * (0.5 - P.beta) * udp + (1.0 - P.gamma) * dt * vpm + pa_in +
                       P.gamma * dt * vppr + P.beta * dt * dt * xbar;
    udb[id] = xbar;
    vb[id] = vppr;
    pab[id] = -P.alpha * ia * app;
    const double abar = free_d ? app * ia : 0.0;
    const double a = ((P.has_fe ? fe[id] : 0.0) - fi[id]) / m;
    double fb = -abar / m;
    if (d == 2 && rgroup[i] >= 0) fb += hbar_k[rgroup[i]];
    fib[id] = fb;
    mbar -= abar * a / m;
    if (dtb) {
      const double ac = ia * ((free_d ? a : 0.0) - P.alpha * parec[id]);
      const double vv = free_d ? vrec[id] + P.gamma * dt * ac : vrec[id];  // an imposed v' is its value
      dtbar += udp * (vv + 2.0 * (0.5 - P.beta) * dt * ac) + vpm * (1.0 - P.gamma) * ac +
               vppr * P.gamma * ac + xbar * 2.0 * P.beta * dt * ac;
    }
  }
  mb[i] = mbar;
  if (dtb) dtb[i] = dtbar;
}

// Per-step CFL time step dt = cfl m / (cs + v), m the smallest edge of the
// first four nodes of every element (calcMinEdgeLength's edges, Domain_d.C:
// 2224-2260) and v the largest speed of the predicted velocities
// (Solver_explicit.C:576-600), both at the step's start.  One block finds an
// element edge of length m and a node of speed v (ties: the lowest index)
// from the step record: loc = {m, 6 e + edge, v, node}.
constexpr int kCflB = 1024;
__device__ const int kEdgePairs[6][2] = {{1, 0}, {2, 0}, {3, 0}, {2, 1}, {3, 1}, {3, 2}};
__global__ __launch_bounds__(kCflB) void a_cfl_locate(int E, int N, const int *elnod, const double *rec,
                                                       double *loc) {
  __shared__ double sv[kCflB];
  __shared__ long si[kCflB];
  const long Ns = N, Es = E;
  const double *x = rec, *v = rec + 3 * Ns;
  const int t = threadIdx.x;
  double m = 1.0e300;
  long mi = -1;
  for (int e = t; e < E; e += kCflB) {
    double P[4][3];
    for (int a = 0; a < 4; ++a) {
      const int g = elnod[a * Es + e];
      for (int c = 0; c < 3; ++c) P[a][c] = x[c * Ns + g];
    }
    for (int q = 0; q < 6; ++q) {
      const int a = kEdgePairs[q][0], b = kEdgePairs[q][1];
      const double ex = P[a][0] - P[b][0], ey = P[a][1] - P[b][1], ez = P[a][2] - P[b][2];
      const double len = sqrt(ex * ex + ey * ey + ez * ez);
      if (len < m) { m = len; mi = 6L * e + q; }
    }
  }
  double vm = -1.0;
  long vi = -1;
  for (int i = t; i < N; i += kCflB) {
    const double vx = v[i], vy = v[Ns + i], vz = v[2 * Ns + i];
    const double sp = sqrt(vx * vx + vy * vy + vz * vz);
    if (sp > vm) { vm = sp; vi = i; }
  }
  // (value, index) extremes over the block, the lower index on ties
  for (int pass = 0; pass < 2; ++pass) {
    const bool mn = pass == 0;
    sv[t] = mn ? m : vm;
    si[t] = mn ? mi : vi;
    __syncthreads();
    for (int h = kCflB / 2; h > 0; h >>= 1) {
      if (t < h) {
        const double a = sv[t], b = sv[t + h];
        const long ia = si[t], ib = si[t + h];
        const bool take = ib >= 0 && (ia < 0 || (mn ? (b < a || (b == a && ib < ia)) : (b > a || (b == a && ib < ia))));
        if (take) { sv[t] = b; si[t] = ib; }
      }
      __syncthreads();
    }
    if (t == 0) {
      loc[mn ? 0 : 2] = sv[0];
      loc[mn ? 1 : 3] = (double)si[0];
    }
    __syncthreads();
  }
}

// The step length's cotangent (node and element shares summed in a fixed
// order) passed on through dt = cfl m / (cs + v): dm-bar = dt-bar dt / m to the
// two nodes of the edge found, dv-bar = -dt-bar dt^2 / (cfl m) to the node found
// (cs + v = cfl m / dt).
__global__ __launch_bounds__(kCflB) void a_cfl_apply(int N, int E, double dt, double cfl, const double *dtb,
                                                      const double *loc, const int *elnod, const double *rec,
                                                      double *xb, double *vb) {
  __shared__ double sh[kCflB];
  const long Ns = N, Es = E;
  const int t = threadIdx.x;
  double s = 0.0;
  for (long k = t; k < Ns + Es; k += kCflB) s += dtb[k];
  sh[t] = s;
  __syncthreads();
  for (int h = kCflB / 2; h > 0; h >>= 1) {
    if (t < h) sh[t] += sh[t + h];
    __syncthreads();
  }
  if (t != 0) return;
  const double dtbar = sh[0], m = loc[0], vmax = loc[2];
  const long mi = (long)loc[1], vi = (long)loc[3];
  if (mi >= 0 && m > 0.0) {
    const long e = mi / 6;
    const int q = (int)(mi % 6);
    const int ga = elnod[kEdgePairs[q][0] * Es + e], gb = elnod[kEdgePairs[q][1] * Es + e];
    const double mbar = dtbar * dt / m;
    for (int c = 0; c < 3; ++c) {
      const double dx = rec[c * Ns + ga] - rec[c * Ns + gb];
      xb[c * Ns + ga] += mbar * dx / m;
      xb[c * Ns + gb] -= mbar * dx / m;
    }
  }
  if (vi >= 0 && vmax > 0.0 && m > 0.0) {
    const double vbar = -dtbar * dt * dt / (cfl * m);
    for (int c = 0; c < 3; ++c) vb[c * Ns + vi] += vbar * rec[3 * Ns + c * Ns + vi] / vmax;
  }
}

// The rate laws' temperature factors of each element at its own temperature
// (per-element temperatures, fixed during a run), as the forward forms them:
// Johnson-Cook 1 - T_h^m (Material.cuh:379-384); GMT at T clamped to
// [T_min, T_max]: C1 exp(C2 T), exp(C2 T), n1 T + n2, m1 T + m2, I1 T + I2
// (Material.cuh:441-444, :465-472).  lawc [5][E].
__global__ void k_law_temps(int E, int law, const double *Te, double T_t, double T_m, double mj, double C1,
                            double C2, double n1, double n2, double m1, double m2, double I1, double I2,
                            double T_min, double T_max, double *lawc) {
  int e = (int)tid();
  if (e >= E) return;
  const double T = Te[e];
  if (law == 2) {
    lawc[e] = 1.0 - pow((T - T_t) / (T_m - T_t), mj);
  } else {
    const double Tc = T < T_min ? T_min : (T > T_max ? T_max : T);
    lawc[e] = C1 * exp(C2 * Tc);
    lawc[E + e] = exp(C2 * Tc);
    lawc[2L * E + e] = n1 * Tc + n2;
    lawc[3L * E + e] = m1 * Tc + m2;
    lawc[4L * E + e] = I1 * Tc + I2;
  }
}

template <int NN>
__global__ __launch_bounds__(kB) void a_elem1(int E, int N, adj::ElemConst K, const int *elnod,
                                              const double *rec, const double *fib, const double *sy0d,
                                              const double *Hd, const double *rhod, double *taub,
                                              double *epb, double *contrib, double *volb, double *grad,
                                              double *dtb_e, const double *lawc) {
  int e = (int)tid();
  if (e >= E) return;
  if (lawc) {  // the element's own temperature factors
    if (K.law == 2) {
      K.jsoft = lawc[e];
    } else {
      K.ga = lawc[e];
      K.gexp = lawc[E + e];
      K.gn = lawc[2L * E + e];
      K.gm = lawc[3L * E + e];
      K.gi = lawc[4L * E + e];
    }
  }
  const long Es = E, Ns = N;
  const double *x = rec, *v = rec + 3 * Ns, *r = rec + 10 * Ns, *tau = rec + 11 * Ns, *ep = tau + 6 * Es;
  adj::ElemIn<NN> in;
  double fb[NN][3];
  for (int a = 0; a < NN; ++a) {
    const int g = elnod[a * Es + e];
    for (int c = 0; c < 3; ++c) {
      in.X[a][c] = x[c * Ns + g];
      in.V[a][c] = v[c * Ns + g];
      fb[a][c] = fib[c * Ns + g];
    }
    in.r[a] = r[g];
  }
  double tb[6];
  for (int q = 0; q < 6; ++q) {
    in.tau[q] = tau[q * Es + e];
    tb[q] = taub[q * Es + e];
  }
  in.ep = ep[e];
  in.rho = rhod[e];
  in.sy0 = sy0d[e];
  in.H = Hd[e];
  adj::ElemFwd<NN> F;
  adj::elem_forward<NN>(K, in, F);
  adj::ElemBar<NN> B;
  adj::elem_reverse<NN>(K, in, F, fb, tb, epb[e], B);
  for (int q = 0; q < 6; ++q) taub[q * Es + e] = B.tau[q];
  epb[e] = B.ep;
  for (int a = 0; a < NN; ++a) {
    for (int c = 0; c < 3; ++c) {
      contrib[(a * 7 + c) * Es + e] = B.X[a][c];
      contrib[(a * 7 + 3 + c) * Es + e] = B.V[a][c];
    }
    contrib[(a * 7 + 6) * Es + e] = B.r[a];
  }
  volb[e] = B.volb;
  if (dtb_e) dtb_e[e] = B.dt;
  // rows: the law's two parameters (bilinear sy0, H; Hollomon Kh, mh;
  // Johnson-Cook A, B; GMT C1, n2), hg
  grad[e] += K.law == 1 ? B.Kh : B.sy0;
  grad[Es + e] += K.law == 1 ? B.mh : B.H;
  grad[2 * Es + e] += B.hg;
}

__global__ void a_node2(int N, int E, int n, const int *off, const int *cnt, const int *pk, const double *contrib,
                        const double *vsum0, const double *mb, const double *rho_avg, double *w) {
  int i = (int)tid();
  if (i >= N) return;
  const long Es = E;
  double rs = 0.0;
  for (int k = 0; k < cnt[i]; ++k) {
    const int q = pk[off[i] + k];
    rs += contrib[((q & 7) * 7 + 6) * Es + (q >> 3)];
  }
  w[i] = rs / vsum0[i] + mb[i] * rho_avg[i] / n;
}

template <int NN>
__global__ __launch_bounds__(kB) void a_elem2(int E, int N, const int *elnod, const double *rec,
                                              const double *volb, const double *w, double *contrib) {
  int e = (int)tid();
  if (e >= E) return;
  const long Es = E, Ns = N;
  double X[NN][3], vb = volb[e];
  for (int a = 0; a < NN; ++a) {
    const int g = elnod[a * Es + e];
    for (int c = 0; c < 3; ++c) X[a][c] = rec[c * Ns + g];
    vb += w[g];
  }
  double Xb[NN][3];
  adj::vol_reverse<NN>(X, vb, Xb);
  for (int a = 0; a < NN; ++a)
    for (int c = 0; c < 3; ++c) contrib[(a * 7 + c) * Es + e] += Xb[a][c];
}

__global__ void a_node3(int N, int E, const int *off, const int *cnt, const int *pk, const double *contrib,
                        double *xb, double *vb) {
  int i = (int)tid();
  if (i >= N) return;
  const long Es = E, Ns = N;
  double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
  for (int k = 0; k < cnt[i]; ++k) {
    const int q = pk[off[i] + k];
    const long base = (long)(q & 7) * 7, e = q >> 3;
    for (int c = 0; c < 3; ++c) {
      sx[c] += contrib[(base + c) * Es + e];
      sv[c] += contrib[(base + 3 + c) * Es + e];
    }
  }
  for (int c = 0; c < 3; ++c) {
    xb[c * Ns + i] += sx[c];
    vb[c * Ns + i] += sv[c];
  }
}

// parameter gradients: one work-group per parameter row, fixed order; for
// the bilinear law the tangent-modulus row is converted from H to Et on the
// way (H = E Et/(E - Et), dH/dEt = (E + H)^2 / E^2)
__global__ void a_reduce(int E, double Emod, int bilinear, const double *grad, const double *Hd, double *gEt,
                         double *out) {
  __shared__ double sh[1024];
  const int row = blockIdx.x;
  double s = 0.0;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    double g = grad[(long)row * E + e];
    if (row == 1) {
      if (bilinear) {
        const double f = (Emod + Hd[e]) / Emod;
        g *= f * f;
      }
      gEt[e] = g;
    }
    s += g;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[row] = sh[0];
}

}  // namespace

hipError_t launch_reaction_f32(wfx_ctx *c, const float *fi32, double *dev_out, hipStream_t s);
std::vector<std::pair<void *, size_t>> fused_state_buffers(wfx_ctx *c);

}  // namespace wfx

using namespace wfx;

#define REQUIRE(cond, code, msg)        \
  do {                                  \
    if (!(cond)) {                      \
      wfx::set_error(msg);              \
      return code;                      \
    }                                   \
  } while (0)
#define LAUNCH(expr) WFX_CHECK_HIP(expr)

static hipStream_t as_stream(wfx_stream s) { return (hipStream_t)s; }

template <typename T>
static hipError_t node_arrays(wfx_ctx *c, T **v, T **ud, T **pa, T **fi) {
  if constexpr (std::is_same<T, double>::value) {
    *v = c->nd.v; *ud = c->nd.u_dt; *pa = c->nd.prev_a; *fi = c->nd.fi;
  } else {
    *v = c->v32; *ud = c->ud32; *pa = c->pa32; *fi = c->fi32;
  }
  return hipSuccess;
}

template <typename T>
static DevElem<T> &el_of(wfx_ctx *c);
template <> DevElem<double> &el_of<double>(wfx_ctx *c) { return c->e64; }
template <> DevElem<float> &el_of<float>(wfx_ctx *c) { return c->e32; }

static int check_supported(wfx_ctx *c) {
  const Stab &s = c->prm.stab;
  REQUIRE(c->prm.mat.model == WFX_BILINEAR || c->prm.mat.model == WFX_HOLLOMON ||
              c->prm.mat.model == WFX_JOHNSON_COOK || c->prm.mat.model == WFX_GMT,
          WFX_ERR_STATE, "the adjoint supports the bilinear, Hollomon, Johnson-Cook and GMT materials");
  // thermal coupling one way: the temperatures follow the mechanics, which do
  // not read them when the law ignores temperature and nothing expands, so
  // the mechanical adjoint is the whole one
  REQUIRE(!c->thermal.on || ((c->prm.mat.model == WFX_BILINEAR || c->prm.mat.model == WFX_HOLLOMON) &&
                             c->thermal.exp_T == 0.0),
          WFX_ERR_STATE,
          "the adjoint covers thermal coupling only with a law that ignores temperature and no thermal expansion");
  REQUIRE(!c->contact.on, WFX_ERR_STATE, "the adjoint does not cover contact");
  REQUIRE(s.alpha_free == 0.0 && s.log_factor == 0.0 && s.J_min == 0.0 && s.hg_coeff_free == 0.0 &&
              s.pspg_scale == 0.0 && s.p_pspg_bulkfac == 0.0 && s.av_coeff_div == 0.0 && s.av_coeff_bulk == 0.0 &&
              s.artifvisc0 == 0.0 && s.artifvisc1 == 0.0,
          WFX_ERR_STATE, "the adjoint supports the ANP pressure without stabilisation terms");
  return WFX_OK;
}

static int adj_alloc(wfx_ctx *c, int nsteps, int every) {
  const long N = c->N, E = c->E, n = c->n;
  const int ng = c->ngroups > 0 ? c->ngroups : 1;
  const int nck = (nsteps + every - 1) / every;
  Adjoint *B = c->adj;
  const bool reuse = B && B->nsteps == nsteps && B->every == every && B->ng == ng && B->N == N && B->E == E &&
                     B->precision == c->precision;  // same run shape (identification loops)
  if (!reuse) {
    adjoint_free(c);
    Adjoint *A = new Adjoint();
    c->adj = A;
    A->nsteps = nsteps; A->every = every; A->nck = nck; A->ng = ng;
    A->N = (int)N; A->E = (int)E; A->n = (int)n; A->precision = c->precision;
    A->ck_stride = 12 * N + 7 * E;
    A->slot_stride = 14 * N + 7 * E;
    hipError_t err = hipSuccess;
    auto get = [&](double **p, size_t count) {
      if (err != hipSuccess) return;
      void *q = nullptr;
      err = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(double));
      if (err == hipSuccess) { A->mem.push_back(q); *p = (double *)q; }
    };
    get(&A->ck, A->ck_stride * nck);
    get(&A->slot, A->slot_stride * every);
    get(&A->hist, (size_t)nsteps * ng);
    get(&A->hbar, (size_t)nsteps * ng);
    get(&A->xb, 3 * N); get(&A->vb, 3 * N); get(&A->udb, 3 * N); get(&A->pab, 3 * N);
    get(&A->taub, 6 * E); get(&A->epb, E); get(&A->fib, 3 * N); get(&A->mb, N); get(&A->w, N);
    get(&A->contrib, 7 * n * E); get(&A->volb, E); get(&A->grad, 3 * E);
    get(&A->sy0d, E); get(&A->Hd, E); get(&A->rhod, E); get(&A->rho_avg, N); get(&A->sums, 8 + E);
    get(&A->dts, nsteps); get(&A->dtb, N + E); get(&A->cfl_loc, 8); get(&A->lawc, 5 * E);
    if (err == hipSuccess) {
      void *q = nullptr;
      err = hipMalloc(&q, std::max<long>(N, 1) * sizeof(int));
      if (err == hipSuccess) { A->mem.push_back(q); A->rgroup = (int *)q; }
    }
    // buffers saved at the end of the forward run and restored after the backward
    if (err == hipSuccess) A->final_bufs = fused_state_buffers(c);
    for (size_t k = 0; k < A->final_bufs.size() && err == hipSuccess; ++k) {
      void *q = nullptr;
      err = hipMalloc(&q, A->final_bufs[k].second);
      if (err == hipSuccess) A->final_saved.push_back(q);
    }
    if (err != hipSuccess) {
      adjoint_free(c);
      WFX_CHECK_HIP(err);
    }
  }
  Adjoint *A = c->adj;
  A->recorded = false;
  // node -> reaction group (each node in at most one group)
  std::vector<int> rg(N, -1);
  for (int g = 0; g + 1 < (int)c->react_off_h.size(); ++g)
    for (int k = c->react_off_h[g]; k < c->react_off_h[g + 1]; ++k) {
      const int node = c->react_h[k];
      if (rg[node] != -1) {
        set_error("the adjoint needs every reaction node in at most one group, once");
        return WFX_ERR_STATE;
      }
      rg[node] = g;
    }
  WFX_CHECK_HIP(hipMemcpy(A->rgroup, rg.data(), N * sizeof(int), hipMemcpyHostToDevice));
  return WFX_OK;
}

template <typename T>
static hipError_t save_ck(wfx_ctx *c, int idx, hipStream_t st) {
  Adjoint *A = c->adj;
  T *v, *ud, *pa, *fi;
  node_arrays<T>(c, &v, &ud, &pa, &fi);
  auto &el = el_of<T>(c);
  const long N = c->N, E = c->E;
  hipLaunchKernelGGL((k_ck_save<T>), dim3(nb(A->ck_stride)), dim3(kB), 0, st, N, E, c->nd.x, v, ud, pa, el.tau,
                     el.pl_strain, A->ck + (size_t)idx * A->ck_stride);
  return hipGetLastError();
}

template <typename T>
static hipError_t load_ck(wfx_ctx *c, int idx, hipStream_t st) {
  Adjoint *A = c->adj;
  T *v, *ud, *pa, *fi;
  node_arrays<T>(c, &v, &ud, &pa, &fi);
  auto &el = el_of<T>(c);
  const long N = c->N, E = c->E;
  hipLaunchKernelGGL((k_ck_load<T>), dim3(nb(A->ck_stride)), dim3(kB), 0, st, N, E,
                     A->ck + (size_t)idx * A->ck_stride, c->nd.x, v, ud, pa, el.tau, el.pl_strain);
  return hipGetLastError();
}

// one fused step (phases 1..4); rec != nullptr records its inputs and node
// values.  As in a fused run, only the run's first step reads the predicted
// increment (the node pass forms it again from v' and prev_a after that), so
// a segment recomputed from its checkpoint repeats the forward bit for bit.
// With the per-step CFL time step (step k >= 0) each step forms its dt on the
// device as a fused run does, and reads and stores the increment every step.
template <typename T>
static hipError_t fused_step(wfx_ctx *c, int predict, int k, bool last, double *rec, hipStream_t st) {
  T *v, *ud, *pa, *fi;
  node_arrays<T>(c, &v, &ud, &pa, &fi);
  auto &el = el_of<T>(c);
  const long N = c->N, E = c->E;
  const bool cfl = c->adj->cfl;
  if (rec) {
    hipLaunchKernelGGL((k_rec_pre<T>), dim3(nb(9 * N + 7 * E)), dim3(kB), 0, st, N, E, c->nd.x, v, el.tau,
                       el.pl_strain, pa, rec);
  }
  c->fused_read_ud = (k == 0 || cfl) ? 1 : 0;
  c->fused_run_first = k == 0 ? 1 : 0;
  c->fused_run_last = last ? 1 : 0;
  c->fused_write_ud = cfl ? 1 : 0;
  hipError_t e = hipSuccess;
  for (int ph = 1; ph <= 4 && e == hipSuccess; ++ph) e = launch_fused_phase<T>(c, ph, predict, st, cfl ? k : -1);
  c->fused_read_ud = c->fused_write_ud = c->fused_run_first = c->fused_run_last = 1;
  if (e != hipSuccess) return e;
  if (rec) {
    hipLaunchKernelGGL((k_rec_post<T>), dim3(nb(5 * N)), dim3(kB), 0, st, N, fi, c->nd.mdiag, c->nd.anp_ratio,
                       rec);
  }
  return hipGetLastError();
}

template <typename T>
static int adjoint_forward(wfx_ctx *c, int nsteps, bool want_hist, hipStream_t st) {
  Adjoint *A = c->adj;
  LAUNCH(launch_vsum0<T>(c, st));
  LAUNCH(launch_update_prediction(c, st));
  for (int d = 0; d < 3; ++d) LAUNCH(launch_impose(c, d, false, st));
  LAUNCH(launch_fused_pack<T>(c, st));
  if (A->cfl) LAUNCH(launch_cfl_begin<T>(c, st, true));
  T *v, *ud, *pa, *fi;
  node_arrays<T>(c, &v, &ud, &pa, &fi);
  for (int k = 0; k < nsteps; ++k) {
    if (k % A->every == 0) LAUNCH(save_ck<T>(c, k / A->every, st));
    LAUNCH(fused_step<T>(c, k + 1 < nsteps, k, k + 1 == nsteps, nullptr, st));
    if (want_hist && c->ngroups > 0) {
      double *h = A->hist + (size_t)k * A->ng;
      if constexpr (std::is_same<T, double>::value) LAUNCH(launch_reaction(c, 2, h, st));
      else LAUNCH(launch_reaction_f32(c, c->fi32, h, st));
    }
  }
  LAUNCH(launch_fused_unpack<T>(c, st));
  LAUNCH(launch_fused_phase<T>(c, 5, 0, st, -1));
  LAUNCH(launch_copy(c->nd.prev_a, c->nd.a, 3L * c->N, st));
  for (size_t k = 0; k < A->final_bufs.size(); ++k)
    WFX_CHECK_HIP(hipMemcpyAsync(A->final_saved[k], A->final_bufs[k].first, A->final_bufs[k].second,
                                 hipMemcpyDeviceToDevice, st));
  return WFX_OK;
}

template <typename T>
static int adjoint_backward(wfx_ctx *c, hipStream_t st) {
  Adjoint *A = c->adj;
  const int N = c->N, E = c->E, n = c->n;
  auto &el = el_of<T>(c);
  const auto &P = c->prm;
  for (double *p : {A->xb, A->vb, A->udb, A->pab, A->fib})
    hipLaunchKernelGGL(k_zero, dim3(nb(3L * N)), dim3(kB), 0, st, 3L * N, p);
  hipLaunchKernelGGL(k_zero, dim3(nb(6L * E)), dim3(kB), 0, st, 6L * E, A->taub);
  hipLaunchKernelGGL(k_zero, dim3(nb(E)), dim3(kB), 0, st, (long)E, A->epb);
  hipLaunchKernelGGL(k_zero, dim3(nb(3L * E)), dim3(kB), 0, st, 3L * E, A->grad);
  hipLaunchKernelGGL((k_prep_elem<T>), dim3(nb(E)), dim3(kB), 0, st, E, el.sy0, el.He, P.mat.H, el.rho, A->sy0d,
                     A->Hd, A->rhod);
  hipLaunchKernelGGL(k_prep_node, dim3(nb(N)), dim3(kB), 0, st, N, c->ms.nodel_offset, c->ms.nodel_count,
                     c->ms.nodel_pk, A->rhod, A->rho_avg);
  LAUNCH(hipGetLastError());
  adj::ElemConst K{P.dt, P.mat.G, P.mat.K, P.mat.cs0, P.stab.hg_coeff, n == 8};
  if (P.mat.model == WFX_HOLLOMON) {
    K.law = 1;
    K.Kh = P.mat.Kh;
    K.mh = P.mat.mh;
    K.eps0 = P.mat.eps0;
    K.eps1 = P.mat.eps1;
    K.sy0m = P.mat.sy0;
  } else if (P.mat.model == WFX_JOHNSON_COOK) {
    K.law = 2;
    K.jB = P.mat.B;
    K.jn = P.mat.n;
    K.jC = P.mat.C;
    K.jeps0 = P.mat.eps_0;
    // the softening factor 1 - T_h^m of the uniform temperature (Material.cuh:379-384)
    K.jsoft = 1.0 - pow((P.mat.temp - P.mat.T_t) / (P.mat.T_m - P.mat.T_t), P.mat.m);
    K.jcap = P.mat.max_edot;
    K.Emod = P.mat.E;
  } else if (P.mat.model == WFX_GMT) {
    K.law = 3;
    // the temperature factors at the clamped uniform temperature (Material.cuh:441-444, :465-472)
    const double Tc = std::min(std::max(P.mat.temp, P.mat.T_min), P.mat.T_max);
    K.gexp = exp(P.mat.C2 * Tc);
    K.ga = P.mat.C1 * exp(P.mat.C2 * Tc);
    K.gn = P.mat.n1 * Tc + P.mat.n2;
    K.gm = P.mat.m1 * Tc + P.mat.m2;
    K.gi = P.mat.I1 * Tc + P.mat.I2;
    K.ge_min = P.mat.e_min;
    K.ge_max = P.mat.e_max;
    K.ger_min = P.mat.er_min;
    K.ger_max = P.mat.er_max;
    K.jcap = P.mat.max_edot;
  }
  // per-element temperatures: the rate laws' factors of every element
  const double *lawc = nullptr;
  if (c->temp_e && (K.law == 2 || K.law == 3)) {
    const Material &M = P.mat;
    hipLaunchKernelGGL(k_law_temps, dim3(nb(E)), dim3(kB), 0, st, E, K.law, c->temp_e, M.T_t, M.T_m, M.m, M.C1,
                       M.C2, M.n1, M.n2, M.m1, M.m2, M.I1, M.I2, M.T_min, M.T_max, A->lawc);
    LAUNCH(hipGetLastError());
    lawc = A->lawc;
  }
  NodeAdj NP{P.dt, P.alpha, P.beta, P.gamma, c->has_fe, A->ng};
  for (int seg = A->nck - 1; seg >= 0; --seg) {
    const int k0 = seg * A->every, k1 = std::min(A->nsteps, k0 + A->every);
    LAUNCH(load_ck<T>(c, seg, st));
    if (A->cfl) LAUNCH(launch_cfl_begin<T>(c, st, true));  // the first recomputed step's largest speed
    for (int k = k0; k < k1; ++k)
      LAUNCH(fused_step<T>(c, k + 1 < A->nsteps, k, k + 1 == A->nsteps, A->slot + (size_t)(k - k0) * A->slot_stride,
                           st));
    for (int k = k1 - 1; k >= k0; --k) {
      const double *rec = A->slot + (size_t)(k - k0) * A->slot_stride;
      if (A->cfl) NP.dt = K.dt = A->dts_h[k];
      double *dtb_n = A->cfl ? A->dtb : nullptr, *dtb_e = A->cfl ? A->dtb + N : nullptr;
      hipLaunchKernelGGL(a_node1, dim3(nb(N)), dim3(kB), 0, st, N, E, NP, rec, c->nd.fe, c->bc_slot, c->bc_mask,
                         A->rgroup, A->hbar + (size_t)k * A->ng, A->xb, A->vb, A->udb, A->pab, A->fib, A->mb,
                         dtb_n);
      if (n == 8)
        hipLaunchKernelGGL((a_elem1<8>), dim3(nb(E)), dim3(kB), 0, st, E, N, K, c->ms.elnod, rec, A->fib,
                           A->sy0d, A->Hd, A->rhod, A->taub, A->epb, A->contrib, A->volb, A->grad, dtb_e, lawc);
      else
        hipLaunchKernelGGL((a_elem1<4>), dim3(nb(E)), dim3(kB), 0, st, E, N, K, c->ms.elnod, rec, A->fib,
                           A->sy0d, A->Hd, A->rhod, A->taub, A->epb, A->contrib, A->volb, A->grad, dtb_e, lawc);
      hipLaunchKernelGGL(a_node2, dim3(nb(N)), dim3(kB), 0, st, N, E, n, c->ms.nodel_offset, c->ms.nodel_count,
                         c->ms.nodel_pk, A->contrib, c->nd.voln0_sum, A->mb, A->rho_avg, A->w);
      if (n == 8)
        hipLaunchKernelGGL((a_elem2<8>), dim3(nb(E)), dim3(kB), 0, st, E, N, c->ms.elnod, rec, A->volb, A->w,
                           A->contrib);
      else
        hipLaunchKernelGGL((a_elem2<4>), dim3(nb(E)), dim3(kB), 0, st, E, N, c->ms.elnod, rec, A->volb, A->w,
                           A->contrib);
      hipLaunchKernelGGL(a_node3, dim3(nb(N)), dim3(kB), 0, st, N, E, c->ms.nodel_offset, c->ms.nodel_count,
                         c->ms.nodel_pk, A->contrib, A->xb, A->vb);
      if (A->cfl) {  // the step length's dependence on x and v' at the step's start
        hipLaunchKernelGGL(a_cfl_locate, dim3(1), dim3(kCflB), 0, st, E, N, c->ms.elnod, rec, A->cfl_loc);
        hipLaunchKernelGGL(a_cfl_apply, dim3(1), dim3(kCflB), 0, st, N, E, A->dts_h[k], P.cfl, A->dtb, A->cfl_loc,
                           c->ms.elnod, rec, A->xb, A->vb);
      }
      LAUNCH(hipGetLastError());
    }
  }
  // the forward run's final state again
  for (size_t k = 0; k < A->final_bufs.size(); ++k)
    WFX_CHECK_HIP(hipMemcpyAsync(A->final_bufs[k].first, A->final_saved[k], A->final_bufs[k].second,
                                 hipMemcpyDeviceToDevice, st));
  LAUNCH(launch_copy(c->nd.prev_a, c->nd.a, 3L * c->N, st));
  hipLaunchKernelGGL(a_reduce, dim3(3), dim3(1024), 0, st, E, P.mat.E, P.mat.model == WFX_BILINEAR ? 1 : 0,
                     A->grad, A->Hd, A->sums + 8, A->sums);
  LAUNCH(hipGetLastError());
  return WFX_OK;
}

int wfx_adjoint_forward(wfx_ctx *c, int nsteps, int checkpoint_every, double *reaction_hist, wfx_stream s) {
  REQUIRE(c, WFX_ERR_ARG, "null context");
  REQUIRE(c->has_mesh && c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(nsteps > 0 && checkpoint_every > 0, WFX_ERR_ARG, "nsteps and checkpoint_every must be positive");
  int rc = check_supported(c);
  if (rc) return rc;
  rc = adj_alloc(c, nsteps, checkpoint_every);
  if (rc) return rc;
  hipStream_t st = as_stream(s);
  Adjoint *A = c->adj;
  A->cfl = c->prm.cfl > 0.0;
  // each step's dt goes to the adjoint's own record (the per-step CFL reduction writes dt_hist[step])
  double *hist_keep = c->dt_hist;
  const int cap_keep = c->dt_cap;
  c->dt_hist = A->dts;
  c->dt_cap = nsteps;
  rc = c->precision == WFX_FP64 ? adjoint_forward<double>(c, nsteps, reaction_hist != nullptr, st)
                                : adjoint_forward<float>(c, nsteps, reaction_hist != nullptr, st);
  c->dt_hist = hist_keep;
  c->dt_cap = cap_keep;
  if (rc) return rc;
  c->fused_stale = true;
  if (A->cfl) {  // the host's dt follows the run's last step, as after a fused run
    A->dts_h.resize(nsteps);
    WFX_CHECK_HIP(hipMemcpyAsync(A->dts_h.data(), A->dts, nsteps * sizeof(double), hipMemcpyDeviceToHost, st));
    WFX_CHECK_HIP(hipStreamSynchronize(st));
    c->prm.dt = A->dts_h.back();
  }
  if (reaction_hist) {
    const size_t bytes = (size_t)nsteps * c->adj->ng * sizeof(double);
    if (c->ngroups > 0) WFX_CHECK_HIP(hipMemcpyAsync(reaction_hist, c->adj->hist, bytes, hipMemcpyDeviceToHost, st));
    else std::fill(reaction_hist, reaction_hist + (size_t)nsteps * c->adj->ng, 0.0);
  }
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  c->adj->recorded = true;
  return WFX_OK;
}

int wfx_adjoint_backward(wfx_ctx *c, const double *reaction_hist_bar, double *grad, double *grad_sy0_elem,
                         double *grad_Et_elem, double *grad_hg_elem, wfx_stream s) {
  REQUIRE(c, WFX_ERR_ARG, "null context");
  REQUIRE(c->adj && c->adj->recorded, WFX_ERR_STATE, "call wfx_adjoint_forward first");
  REQUIRE(reaction_hist_bar && grad, WFX_ERR_ARG, "null argument");
  Adjoint *A = c->adj;
  REQUIRE(A->N == c->N && A->E == c->E && A->precision == c->precision, WFX_ERR_STATE,
          "mesh changed since wfx_adjoint_forward");
  REQUIRE(c->ngroups > 0, WFX_ERR_STATE, "no reaction nodes: the loss has no input");
  REQUIRE(A->ng == c->ngroups, WFX_ERR_STATE, "reaction groups changed since wfx_adjoint_forward");
  REQUIRE(A->cfl == (c->prm.cfl > 0.0), WFX_ERR_STATE, "the time-step rule changed since wfx_adjoint_forward");
  int rc = check_supported(c);
  if (rc) return rc;
  hipStream_t st = as_stream(s);
  WFX_CHECK_HIP(hipMemcpyAsync(A->hbar, reaction_hist_bar, (size_t)A->nsteps * A->ng * sizeof(double),
                               hipMemcpyHostToDevice, st));
  double *hist_keep = c->dt_hist;
  const int cap_keep = c->dt_cap;
  c->dt_hist = A->dts;
  c->dt_cap = A->nsteps;
  const double dt_keep = c->prm.dt;
  rc = c->precision == WFX_FP64 ? adjoint_backward<double>(c, st) : adjoint_backward<float>(c, st);
  c->dt_hist = hist_keep;
  c->dt_cap = cap_keep;
  c->prm.dt = dt_keep;
  if (rc) return rc;
  WFX_CHECK_HIP(hipMemcpyAsync(grad, A->sums, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  // per-element gradients: device rows in the stored (cluster) order, handed
  // back in the original element order
  const size_t E = (size_t)c->E, eb = E * sizeof(double);
  double *outs[3] = {grad_sy0_elem, grad_Et_elem, grad_hg_elem};
  const double *srcs[3] = {A->grad, A->sums + 8, A->grad + 2 * E};
  std::vector<double> rows[3];
  for (int k = 0; k < 3; ++k)
    if (outs[k]) {
      rows[k].resize(E);
      WFX_CHECK_HIP(hipMemcpyAsync(rows[k].data(), srcs[k], eb, hipMemcpyDeviceToHost, st));
    }
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 3; ++k)
    if (outs[k])
      for (size_t e = 0; e < E; ++e) outs[k][e] = rows[k][c->el_o2s[e]];
  return WFX_OK;
}

int wfx_adjoint_release(wfx_ctx *c) {
  REQUIRE(c, WFX_ERR_ARG, "null context");
  adjoint_free(c);
  return WFX_OK;
}
