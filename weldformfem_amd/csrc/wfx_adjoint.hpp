// Element-level forward and reverse (vector-Jacobian product) of the explicit
// step, for the hand-written adjoint.  One element, FP64, no memory traffic:
// the kernels in wfx_adjoint.hip gather the inputs, call these functions and
// scatter the results.  The functions are __host__ __device__ so the same
// source can be checked on the CPU.
//
// Forward map of one element for one step (the per-step sequence of
// Domain_d, see Solver_explicit.C:524-978):
//   J       = js * sum_a S(k,a) X(a,c)                 calcElemJAndDerivatives
//             (hexa: Domain_d.C:1822-1831, js = 1/8; tet: :1855-1887, js = 1)
//   A       = adj(J) (transposed cofactor), detJ = det J (Matrix.h:593-727)
//   dH(c,a) = js * sum_k A(c,k) S(k,a)                 (Domain_d.C:1843-1850)
//   vol     = w * detJ, w = 8 (hexa) or 1/6 (tet)      CalcElemVol (Mechanical.C:264-293)
//   L(c,d)  = (1/detJ) sum_a dH(d,a) V(a,c)            calcElemStrainRates (Mechanical.C:41-126)
//   D = sym L, W = skew L (xy, yz, xz)
//   p       = -K (mean_a r_a - 1), r = nodal V/V0      calcElemPressure, algorithm 0 with
//                                                      zero stabilisation (Mechanical.C:691-819)
//   tau1    = tau + dt (2G dev D + R(tau, W))          CalcStressStrain (Mechanical.C:1690-1698)
//             R is tau*W^T + W*tau with the reference's tensor3 product (x column
//             A*B, y and z columns A*B^T, Tensor3.C), which reduces to
//             R_xx = 2(Wxy txy + Wxz txz), R_yy = R_zz = 0,
//             R_xy = txx Wxy - txz Wyz + Wxy tyy + Wxz tyz,
//             R_yz = txy Wxz + tyy Wyz - Wxy txz + Wyz tzz,
//             R_xz = txx Wxz + txy Wyz + Wxy tyz + Wxz tzz
//   s       = dev(tau1 - p I), st = sqrt(3 J2(s))      (Mechanical.C:1710-1721)
//   sy, H   = sy0 + H ep, H (bilinear); Kh s^mh, Kh mh s^(mh-1) with s = ep + eps0
//             if s > eps1, else sy0 and 0 (Hollomon, Material.cuh:357-360, :391-392);
//             Johnson-Cook (Material.cuh:378-411, rate and cap Mechanical.C:1700-1748):
//             (sy0 + B ep^n)(1 + C ln(r/eps_0)) soft with r the von Mises
//             rate of D (1e-5 at 0), tangent n B ep^(n-1)(1 + C ln(min(r, cap)/eps_0))
//             soft for ep > 0, else 0.1 E; GMT (Material.cuh:412-473):
//             C1 exp(C2 T) es^n exp(I/es) er^m with es, er the clamped plastic
//             strain and rate, tangent C1 exp(C2 T) erc^m es^(n-2) (n es - I) exp(I/es)
//             at the capped rate
//   plastic (sy < st): dg = (st - sy)/(3G + H), tau2 = (1 - 3G dg/st) s, ep2 = ep + dg
//   elastic:           tau2 = tau1, ep2 = ep           (Mechanical.C:1723-1762)
//   sigma   = tau2 - p I                               (:1766)
//   f(a,c)  = w sum_d dH(d,a) sigma(c,d)               calcElemForces (Mechanical.C:448-457)
//   hexa hourglass: hm(c,j) = sum_a V(a,c) G(j,a); fhg(a,c) = -c_h sum_j hm(c,j) G(j,a),
//             c_h = hg vol^0.6666666 rho 0.25 cs0      (f90_ver/src/Mechanical.f90:241-303)
//   element contribution to the internal force: fnet = f - fhg (assemblyForces,
//   Matrices.C:42-87)
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define WFX_HD __host__ __device__ __forceinline__

namespace wfx {
namespace adj {

struct ElemConst {
  double dt, G, K, cs0, hg;
  int hourglass;  // hexa viscous hourglass on
  // hardening law: 0 bilinear (per-element sy0 and H of ElemIn), 1 Hollomon
  // sigma_y = Kh (ep + eps0)^mh above eps1, the plateau sy0m below
  // (Material.cuh:357-360, :391-392)
  int law = 0;
  double Kh = 0.0, mh = 0.0, eps0 = 0.0, eps1 = 0.0, sy0m = 0.0;
  // law 2, Johnson-Cook at a uniform temperature (Material.cuh:378-411):
  // sigma_y = (A_e + B ep^n)(1 + C ln(r/eps_0)) soft, r the effective strain
  // rate (1e-5 for a zero rate); tangent n B ep^(n-1)(1 + C ln(rc/eps_0)) soft
  // with the capped rate rc = min(r, cap) for ep > 0, 0.1 E at ep = 0
  double jB = 0.0, jn = 0.0, jC = 0.0, jeps0 = 1.0, jsoft = 1.0, jcap = 1.0e6, Emod = 0.0;
  // law 3, GMT at a uniform temperature Tc (Material.cuh:412-473):
  // sigma_y = ga es^gn exp(gi/es) er^gm with ga = C1 exp(C2 Tc) = C1 gexp,
  // gn = n1 Tc + n2, gm = m1 Tc + m2, gi = I1 Tc + I2, es and er the plastic
  // strain and rate clamped to [e_min, e_max] and [er_min, er_max]; tangent
  // ga erc^gm es^(gn-2) (gn es - gi) exp(gi/es) at the capped rate (the cap
  // jcap)
  double ga = 0.0, gexp = 0.0, gn = 0.0, gm = 0.0, gi = 0.0;
  double ge_min = 0.0, ge_max = 0.0, ger_min = 0.0, ger_max = 0.0;
};

template <int NN>
struct ElemIn {
  double X[NN][3], V[NN][3], r[NN];
  double tau[6];  // xx yy zz xy yz xz
  double ep, rho, sy0, H;
};

template <int NN>
struct ElemFwd {
  // geometry
  double J[3][3], C[3][3], detJ, dH[3][NN], vol;
  // kinematics
  double Gm[3][3];  // sum_a dH(d,a) V(a,c), stored Gm[c][d]
  double f, D[6], W[3];
  // stress
  double p, t1[6], s[6], st, sy, dg, fac, t2[6], ep2, sig[6];
  double trate[6];  // 2G dev D + R(tau, W): tau1 = tau + dt trate
  double Ht;  // tangent modulus of the return (bilinear H; Hollomon Kh mh s^(mh-1) or 0)
  double hs;  // Hollomon: s = ep + eps0
  int plastic, hpow;  // hpow: Hollomon power branch (s > eps1)
  // Johnson-Cook: effective rate, the rate the yield stress takes, its log
  // term, ep^n, the capped rate and the tangent's log term
  double edot, sr, lterm, pw, edc, ltc;
  // GMT: clamped strain, rate and capped rate, er^gm, erc^gm, exp(gi/es)
  double ges, ger, gerc, gpw, gpwc, gex;
  // forces
  double hm[3][4], ch0, ch, fnet[NN][3];
};

template <int NN> struct Shape;
template <> struct Shape<8> {
  static constexpr double js = 0.125, w = 8.0;
  WFX_HD static double S(int k, int a) {
    // natural-derivative signs of node a: xi, eta, zeta rows
    const int sx = ((a + 1) >> 1) & 1 ? 1 : -1;    // -1 1 1 -1 -1 1 1 -1
    const int sy = (a & 2) ? 1 : -1;                // -1 -1 1 1 -1 -1 1 1
    const int sz = (a & 4) ? 1 : -1;                // -1 -1 -1 -1 1 1 1 1
    return (double)(k == 0 ? sx : (k == 1 ? sy : sz));
  }
};
template <> struct Shape<4> {
  static constexpr double js = 1.0, w = 1.0 / 6.0;
  WFX_HD static double S(int k, int a) { return a == 0 ? -1.0 : (a == k + 1 ? 1.0 : 0.0); }
};

// hexa hourglass base vectors (Flanagan-Belytschko, scaled to +-1)
WFX_HD double hg_gamma(int j, int a) {
  // rows: {1,1,-1,-1,-1,-1,1,1}, {1,-1,-1,1,-1,1,1,-1}, {1,-1,1,-1,1,-1,1,-1}, {-1,1,-1,1,1,-1,1,-1}
  // as bit masks, bit a set for +1
  const unsigned rows[4] = {0xC3u, 0x69u, 0x55u, 0x5Au};
  return ((rows[j] >> a) & 1u) ? 1.0 : -1.0;
}

WFX_HD void cofactor(const double J[3][3], double C[3][3]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      C[r][c] = J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1];
    }
}

template <int NN>
WFX_HD void elem_forward(const ElemConst &k, const ElemIn<NN> &in, ElemFwd<NN> &F) {
  using Sh = Shape<NN>;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int a = 0; a < NN; ++a) s += Sh::S(r, a) * in.X[a][c];
      F.J[r][c] = Sh::js * s;
    }
  cofactor(F.J, F.C);
  F.detJ = F.J[0][0] * F.C[0][0] + F.J[0][1] * F.C[0][1] + F.J[0][2] * F.C[0][2];
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < NN; ++a) {
      double s = 0.0;
      for (int q = 0; q < 3; ++q) s += F.C[q][c] * Sh::S(q, a);  // A(c,q) = C(q,c)
      F.dH[c][a] = Sh::js * s;
    }
  F.vol = Sh::w * F.detJ;
  // velocity gradient
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int a = 0; a < NN; ++a) s += F.dH[d][a] * in.V[a][c];
      F.Gm[c][d] = s;
    }
  F.f = 1.0 / F.detJ;
  double L[3][3];
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) L[c][d] = F.f * F.Gm[c][d];
  F.D[0] = L[0][0]; F.D[1] = L[1][1]; F.D[2] = L[2][2];
  F.D[3] = 0.5 * (L[0][1] + L[1][0]);
  F.D[4] = 0.5 * (L[1][2] + L[2][1]);
  F.D[5] = 0.5 * (L[0][2] + L[2][0]);
  F.W[0] = 0.5 * (L[0][1] - L[1][0]);
  F.W[1] = 0.5 * (L[1][2] - L[2][1]);
  F.W[2] = 0.5 * (L[0][2] - L[2][0]);
  // pressure
  double rs = 0.0;
  for (int a = 0; a < NN; ++a) rs += in.r[a];
  F.p = -k.K * (rs / NN - 1.0);
  // Jaumann update
  const double *t = in.tau;
  const double Wxy = F.W[0], Wyz = F.W[1], Wxz = F.W[2];
  const double R[6] = {2.0 * (Wxy * t[3] + Wxz * t[5]), 0.0, 0.0,
                       t[0] * Wxy - t[5] * Wyz + Wxy * t[1] + Wxz * t[4],
                       t[3] * Wxz + t[1] * Wyz - Wxy * t[5] + Wyz * t[2],
                       t[0] * Wxz + t[3] * Wyz + Wxy * t[4] + Wxz * t[2]};
  const double trD = (F.D[0] + F.D[1] + F.D[2]) / 3.0;
  for (int q = 0; q < 6; ++q) {
    const double dd = q < 3 ? F.D[q] - trD : F.D[q];
    F.trate[q] = 2.0 * k.G * dd + R[q];
    F.t1[q] = t[q] + k.dt * F.trate[q];
  }
  // trial state and radial return
  const double m = ((F.t1[0] - F.p) + (F.t1[1] - F.p) + (F.t1[2] - F.p)) / 3.0;
  for (int q = 0; q < 3; ++q) F.s[q] = (F.t1[q] - F.p) - m;
  for (int q = 3; q < 6; ++q) F.s[q] = F.t1[q];
  const double J2 = 0.5 * (F.s[0] * F.s[0] + F.s[1] * F.s[1] + F.s[2] * F.s[2]) +
                    (F.s[3] * F.s[3] + F.s[4] * F.s[4] + F.s[5] * F.s[5]);
  F.st = sqrt(3.0 * J2);
  if (k.law == 1) {
    F.hs = in.ep + k.eps0;
    F.hpow = F.hs > k.eps1;
    F.sy = F.hpow ? k.Kh * pow(F.hs, k.mh) : k.sy0m;
    F.Ht = F.hpow ? k.Kh * k.mh * pow(F.hs, k.mh - 1.0) : 0.0;
  } else if (k.law == 2) {
    F.hs = 0.0;
    F.hpow = 0;
    const double *D = F.D;
    F.edot = sqrt(0.5 * ((D[0] - D[1]) * (D[0] - D[1]) + (D[1] - D[2]) * (D[1] - D[2]) + (D[2] - D[0]) * (D[2] - D[0])) +
                  3.0 * (D[3] * D[3] + D[4] * D[4] + D[5] * D[5]));
    F.sr = (F.edot == 0.0) ? 1.0e-5 : F.edot;
    F.lterm = 1.0 + k.jC * log(F.sr / k.jeps0);
    F.pw = pow(in.ep, k.jn);
    F.sy = (in.sy0 + k.jB * F.pw) * F.lterm * k.jsoft;
    F.edc = (k.jcap < F.edot) ? k.jcap : F.edot;
    if (in.ep > 0.0) {
      F.ltc = (F.edc == F.sr) ? F.lterm : 1.0 + k.jC * log(F.edc / k.jeps0);
      F.Ht = k.jn * k.jB * pow(in.ep, k.jn - 1.0) * F.ltc * k.jsoft;
    } else {
      F.ltc = 0.0;
      F.Ht = k.Emod * 0.1;
    }
  } else if (k.law == 3) {
    F.hs = 0.0;
    F.hpow = 0;
    const double *D = F.D;
    F.edot = sqrt(0.5 * ((D[0] - D[1]) * (D[0] - D[1]) + (D[1] - D[2]) * (D[1] - D[2]) + (D[2] - D[0]) * (D[2] - D[0])) +
                  3.0 * (D[3] * D[3] + D[4] * D[4] + D[5] * D[5]));
    F.ges = in.ep < k.ge_min ? k.ge_min : (in.ep > k.ge_max ? k.ge_max : in.ep);
    F.ger = F.edot < k.ger_min ? k.ger_min : (F.edot > k.ger_max ? k.ger_max : F.edot);
    F.gpw = pow(F.ger, k.gm);
    F.gex = exp(k.gi / F.ges);
    F.sy = k.ga * pow(F.ges, k.gn) * F.gex * F.gpw;
    F.edc = (k.jcap < F.edot) ? k.jcap : F.edot;
    F.gerc = F.edc < k.ger_min ? k.ger_min : (F.edc > k.ger_max ? k.ger_max : F.edc);
    F.gpwc = (F.gerc == F.ger) ? F.gpw : pow(F.gerc, k.gm);
    F.Ht = k.ga * F.gpwc * pow(F.ges, k.gn - 2.0) * (-k.gi + F.ges * k.gn) * F.gex;
  } else {
    F.hs = 0.0;
    F.hpow = 0;
    F.sy = in.sy0 + in.H * in.ep;
    F.Ht = in.H;
  }
  F.plastic = F.sy < F.st;
  if (F.plastic) {
    F.dg = (F.st - F.sy) / (3.0 * k.G + F.Ht);
    F.fac = 1.0 - (3.0 * k.G * F.dg) / F.st;
    for (int q = 0; q < 6; ++q) F.t2[q] = F.fac * F.s[q];
    F.ep2 = in.ep + F.dg;
  } else {
    F.dg = 0.0;
    F.fac = 1.0;
    for (int q = 0; q < 6; ++q) F.t2[q] = F.t1[q];
    F.ep2 = in.ep;
  }
  for (int q = 0; q < 6; ++q) F.sig[q] = q < 3 ? F.t2[q] - F.p : F.t2[q];
  // element forces
  const double *sg = F.sig;
  for (int a = 0; a < NN; ++a) {
    const double h0 = F.dH[0][a], h1 = F.dH[1][a], h2 = F.dH[2][a];
    F.fnet[a][0] = Sh::w * (h0 * sg[0] + h1 * sg[3] + h2 * sg[5]);
    F.fnet[a][1] = Sh::w * (h1 * sg[1] + h0 * sg[3] + h2 * sg[4]);
    F.fnet[a][2] = Sh::w * (h2 * sg[2] + h1 * sg[4] + h0 * sg[5]);
  }
  F.ch0 = 0.0;
  F.ch = 0.0;
  if (NN == 8 && k.hourglass) {
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < 4; ++j) {
        double s = 0.0;
        for (int a = 0; a < NN; ++a) s += in.V[a][c] * hg_gamma(j, a);
        F.hm[c][j] = s;
      }
    F.ch0 = pow(F.vol, 0.6666666) * in.rho * 0.25 * k.cs0;
    F.ch = k.hg * F.ch0;
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) {
        double g = 0.0;
        for (int j = 0; j < 4; ++j) g -= F.hm[c][j] * hg_gamma(j, a);
        F.fnet[a][c] -= F.ch * g;  // assembly subtracts the hourglass force
      }
  }
}

// Reverse of elem_forward.  Inputs: the forward record F, the cotangents of
// the element's net nodal forces (fb, equal to the internal-force cotangent
// of each node), of the stored deviatoric stress (taub) and of the plastic
// strain (epb).  Outputs: cotangents of X, V, the nodal ratios r, tau, ep,
// the element volume through the hourglass coefficient (volb, the volume
// paths through the nodal volumes are added by the caller) and of the
// parameters: sy0 and H (bilinear), Kh and mh (Hollomon) or A (in sy0) and B
// (in H) (Johnson-Cook), and hg.
template <int NN>
struct ElemBar {
  double X[NN][3], V[NN][3], r[NN], tau[6], ep, volb, sy0, H, Kh, mh, hg;
  double dt;  // the step length's cotangent (through the Jaumann update)
};

template <int NN>
WFX_HD void elem_reverse(const ElemConst &k, const ElemIn<NN> &in, const ElemFwd<NN> &F,
                         const double fb[NN][3], const double taub[6], double epb, ElemBar<NN> &B) {
  using Sh = Shape<NN>;
  double dHb[3][NN];
  for (int a = 0; a < NN; ++a)
    for (int c = 0; c < 3; ++c) {
      B.X[a][c] = 0.0;
      B.V[a][c] = 0.0;
      dHb[c][a] = 0.0;
    }
  B.volb = 0.0;
  B.hg = 0.0;
  // hourglass: fnet -= ch * g(V)
  if (NN == 8 && k.hourglass) {
    double chb = 0.0, hmb[3][4];
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < 4; ++j) hmb[c][j] = 0.0;
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) {
        double g = 0.0;
        for (int j = 0; j < 4; ++j) g -= F.hm[c][j] * hg_gamma(j, a);
        chb -= fb[a][c] * g;
        const double gb = -F.ch * fb[a][c];
        for (int j = 0; j < 4; ++j) hmb[c][j] -= gb * hg_gamma(j, a);
      }
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int j = 0; j < 4; ++j) s += hmb[c][j] * hg_gamma(j, a);
        B.V[a][c] += s;
      }
    B.hg = chb * F.ch0;
    B.volb = chb * k.hg * F.ch0 * 0.6666666 / F.vol;
  }
  // element forces
  double sigb[6] = {0, 0, 0, 0, 0, 0};
  const double *sg = F.sig;
  for (int a = 0; a < NN; ++a) {
    const double h0 = F.dH[0][a], h1 = F.dH[1][a], h2 = F.dH[2][a];
    const double b0 = Sh::w * fb[a][0], b1 = Sh::w * fb[a][1], b2 = Sh::w * fb[a][2];
    sigb[0] += b0 * h0;
    sigb[1] += b1 * h1;
    sigb[2] += b2 * h2;
    sigb[3] += b0 * h1 + b1 * h0;
    sigb[4] += b1 * h2 + b2 * h1;
    sigb[5] += b0 * h2 + b2 * h0;
    dHb[0][a] += b0 * sg[0] + b1 * sg[3] + b2 * sg[5];
    dHb[1][a] += b0 * sg[3] + b1 * sg[1] + b2 * sg[4];
    dHb[2][a] += b0 * sg[5] + b1 * sg[4] + b2 * sg[2];
  }
  // sigma = tau2 - p I
  double t2b[6];
  for (int q = 0; q < 6; ++q) t2b[q] = taub[q] + sigb[q];
  const double pb = -(sigb[0] + sigb[1] + sigb[2]);
  // radial return
  double t1b[6];
  B.ep = epb;
  B.sy0 = 0.0;
  B.H = 0.0;
  B.Kh = 0.0;
  B.mh = 0.0;
  double edb = 0.0;  // Johnson-Cook: cotangent of the effective strain rate
  if (F.plastic) {
    const double G3 = 3.0 * k.G, den = G3 + F.Ht;
    double sb[6], facb = 0.0;
    for (int q = 0; q < 6; ++q) {
      sb[q] = F.fac * t2b[q];
      facb += t2b[q] * F.s[q];
    }
    const double dgb = epb - G3 / F.st * facb;
    double stb = G3 * F.dg / (F.st * F.st) * facb + dgb / den;
    const double syb = -dgb / den;
    const double Htb = -dgb * F.dg / den;
    if (k.law == 1) {
      if (F.hpow) {  // sy = Kh s^mh, Ht = Kh mh s^(mh-1), s = ep + eps0
        const double pm1 = pow(F.hs, k.mh - 1.0), lns = log(F.hs);
        B.ep += syb * F.Ht + Htb * k.Kh * k.mh * (k.mh - 1.0) * pm1 / F.hs;
        B.Kh = syb * pm1 * F.hs + Htb * k.mh * pm1;
        B.mh = syb * F.sy * lns + Htb * k.Kh * pm1 * (1.0 + k.mh * lns);
      }  // plateau: sy and Ht do not depend on ep, Kh or mh
    } else if (k.law == 2) {
      // sy = (A + B ep^n) lterm soft: rows A (sy0) and B (H)
      const double hard = in.sy0 + k.jB * F.pw;
      B.sy0 = syb * F.lterm * k.jsoft;
      B.H = syb * F.pw * F.lterm * k.jsoft;
      if (F.edot != 0.0) edb += syb * hard * k.jsoft * k.jC / F.edot;  // d lterm / d r = C / r
      if (in.ep > 0.0) {
        const double pnm1 = pow(in.ep, k.jn - 1.0);
        B.ep += syb * k.jB * k.jn * pnm1 * F.lterm * k.jsoft;
        // Ht = n B ep^(n-1) ltc soft
        B.ep += Htb * k.jn * (k.jn - 1.0) * k.jB * pow(in.ep, k.jn - 2.0) * F.ltc * k.jsoft;
        B.H += Htb * k.jn * pnm1 * F.ltc * k.jsoft;
        // the capped rate follows the rate below the cap (a zero rate leaves
        // the tangent at ln 0: no derivative)
        if (F.edc == F.edot && F.edot != 0.0) edb += Htb * k.jn * k.jB * pnm1 * k.jsoft * k.jC / F.edot;
      }  // ep = 0: Ht = 0.1 E
    } else if (k.law == 3) {
      // rows C1 (sy0) and n2 (H); the clamps pass nothing outside their ranges
      const double es = F.ges, q = k.gn * es - k.gi;
      const double pn = pow(es, k.gn), pn2 = pow(es, k.gn - 2.0), les = log(es);
      const double hq = F.gpwc * F.gex;  // Ht = ga hq pn2 q
      B.sy0 = syb * k.gexp * pn * F.gex * F.gpw + Htb * k.gexp * hq * pn2 * q;
      B.H = syb * F.sy * les + Htb * k.ga * hq * pn2 * (les * q + es);
      if (in.ep > k.ge_min && in.ep < k.ge_max) {
        // d sy / d es = sy q / es^2;  d Ht / d es = ga hq es^(gn-3) ((gn-2) q + gn es - q gi / es)
        B.ep += syb * F.sy * q / (es * es) +
                Htb * k.ga * hq * pow(es, k.gn - 3.0) * ((k.gn - 2.0) * q + k.gn * es - q * k.gi / es);
      }
      if (F.edot > k.ger_min && F.edot < k.ger_max) edb += syb * F.sy * k.gm / F.ger;
      if (F.edc == F.edot && F.edc > k.ger_min && F.edc < k.ger_max) edb += Htb * F.Ht * k.gm / F.gerc;
    } else {
      B.H = Htb + syb * in.ep;
      B.sy0 = syb;
      B.ep += syb * in.H;
    }
    // st = sqrt(3 J2)
    const double J2b = 1.5 * stb / F.st;
    for (int q = 0; q < 3; ++q) sb[q] += J2b * F.s[q];
    for (int q = 3; q < 6; ++q) sb[q] += 2.0 * J2b * F.s[q];
    // s = dev(tau1 - p I): the pressure cancels
    const double mb = (sb[0] + sb[1] + sb[2]) / 3.0;
    for (int q = 0; q < 3; ++q) t1b[q] = sb[q] - mb;
    for (int q = 3; q < 6; ++q) t1b[q] = sb[q];
  } else {
    for (int q = 0; q < 6; ++q) t1b[q] = t2b[q];
  }
  // Jaumann update
  const double *t = in.tau;
  const double Wxy = F.W[0], Wyz = F.W[1], Wxz = F.W[2];
  double tb[6], Db[6], Wb[3] = {0, 0, 0};
  for (int q = 0; q < 6; ++q) tb[q] = t1b[q];
  B.dt = 0.0;
  for (int q = 0; q < 6; ++q) B.dt += t1b[q] * F.trate[q];
  {
    const double c2 = 2.0 * k.G * k.dt;
    const double mb = (t1b[0] + t1b[1] + t1b[2]) / 3.0;
    for (int q = 0; q < 3; ++q) Db[q] = c2 * (t1b[q] - mb);
    for (int q = 3; q < 6; ++q) Db[q] = c2 * t1b[q];
    if (edb != 0.0 && F.edot > 0.0) {  // r = sqrt(Q(D)), dr/dD = (dQ/dD) / (2 r)
      const double ie = 0.5 / F.edot, *D = F.D;
      Db[0] += edb * ie * (2.0 * D[0] - D[1] - D[2]);
      Db[1] += edb * ie * (2.0 * D[1] - D[0] - D[2]);
      Db[2] += edb * ie * (2.0 * D[2] - D[0] - D[1]);
      Db[3] += edb * ie * 6.0 * D[3];
      Db[4] += edb * ie * 6.0 * D[4];
      Db[5] += edb * ie * 6.0 * D[5];
    }
    const double Rxx = k.dt * t1b[0], Rxy = k.dt * t1b[3], Ryz = k.dt * t1b[4], Rxz = k.dt * t1b[5];
    // R_xx = 2(Wxy txy + Wxz txz)
    Wb[0] += 2.0 * Rxx * t[3]; tb[3] += 2.0 * Rxx * Wxy;
    Wb[2] += 2.0 * Rxx * t[5]; tb[5] += 2.0 * Rxx * Wxz;
    // R_xy = txx Wxy - txz Wyz + Wxy tyy + Wxz tyz
    tb[0] += Rxy * Wxy; Wb[0] += Rxy * (t[0] + t[1]);
    tb[5] -= Rxy * Wyz; Wb[1] -= Rxy * t[5];
    tb[1] += Rxy * Wxy;
    Wb[2] += Rxy * t[4]; tb[4] += Rxy * Wxz;
    // R_yz = txy Wxz + tyy Wyz - Wxy txz + Wyz tzz
    tb[3] += Ryz * Wxz; Wb[2] += Ryz * t[3];
    tb[1] += Ryz * Wyz; Wb[1] += Ryz * (t[1] + t[2]);
    Wb[0] -= Ryz * t[5]; tb[5] -= Ryz * Wxy;
    tb[2] += Ryz * Wyz;
    // R_xz = txx Wxz + txy Wyz + Wxy tyz + Wxz tzz
    tb[0] += Rxz * Wxz; Wb[2] += Rxz * (t[0] + t[2]);
    tb[3] += Rxz * Wyz; Wb[1] += Rxz * t[3];
    Wb[0] += Rxz * t[4]; tb[4] += Rxz * Wxy;
    tb[2] += Rxz * Wxz;
  }
  for (int q = 0; q < 6; ++q) B.tau[q] = tb[q];
  // pressure
  for (int a = 0; a < NN; ++a) B.r[a] = -k.K / NN * pb;
  // velocity gradient
  double Lb[3][3];
  Lb[0][0] = Db[0]; Lb[1][1] = Db[1]; Lb[2][2] = Db[2];
  Lb[0][1] = 0.5 * (Db[3] + Wb[0]); Lb[1][0] = 0.5 * (Db[3] - Wb[0]);
  Lb[1][2] = 0.5 * (Db[4] + Wb[1]); Lb[2][1] = 0.5 * (Db[4] - Wb[1]);
  Lb[0][2] = 0.5 * (Db[5] + Wb[2]); Lb[2][0] = 0.5 * (Db[5] - Wb[2]);
  double fsb = 0.0;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      fsb += Lb[c][d] * F.Gm[c][d];
      const double gb = F.f * Lb[c][d];
      for (int a = 0; a < NN; ++a) {
        B.V[a][c] += gb * F.dH[d][a];
        dHb[d][a] += gb * in.V[a][c];
      }
    }
  double detJb = -fsb * F.f * F.f;
  // dH = js * A S, A(c,q) = C(q,c)
  double Cb[3][3];
  for (int q = 0; q < 3; ++q)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int a = 0; a < NN; ++a) s += dHb[c][a] * Sh::S(q, a);
      Cb[q][c] = Sh::js * s;
    }
  // detJ = sum_c J(0,c) C(0,c) (= det J for every row)
  double Jb[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Jb[r][c] = detJb * F.C[r][c];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      const double b = Cb[r][c];
      Jb[r1][c1] += b * F.J[r2][c2];
      Jb[r2][c2] += b * F.J[r1][c1];
      Jb[r1][c2] -= b * F.J[r2][c1];
      Jb[r2][c1] -= b * F.J[r1][c2];
    }
  for (int a = 0; a < NN; ++a)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int r = 0; r < 3; ++r) s += Jb[r][c] * Sh::S(r, a);
      B.X[a][c] += Sh::js * s;
    }
}

// Cotangent of X through the element volume alone: vol = w det J(X).
template <int NN>
WFX_HD void vol_reverse(const double X[NN][3], double volb, double Xb[NN][3]) {
  using Sh = Shape<NN>;
  double J[3][3], C[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int a = 0; a < NN; ++a) s += Sh::S(r, a) * X[a][c];
      J[r][c] = Sh::js * s;
    }
  cofactor(J, C);
  const double detJb = Sh::w * volb;
  for (int a = 0; a < NN; ++a)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int r = 0; r < 3; ++r) s += C[r][c] * Sh::S(r, a);
      Xb[a][c] = Sh::js * detJb * s;
    }
}

}  // namespace adj
}  // namespace wfx
