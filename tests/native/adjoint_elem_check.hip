// Host-only test program for the element forward/reverse functions of the adjoint
// (weldformfem_amd/csrc/wfx_adjoint.hpp).  Built with hipcc --offload-host-only
// by tests/test_adjoint_math.py; runs on the CPU.
//
//   adjoint_elem_check forward   < elements.txt   element forward, one line per element
//   adjoint_elem_check vjp NN SEED COUNT PLASTIC [LAW]  reverse vs central
//                                    differences (LAW 0 bilinear, 1 Hollomon,
//                                    2 Johnson-Cook, 3 GMT)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../weldformfem_amd/csrc/wfx_adjoint.hpp"

using namespace wfx::adj;

static unsigned long long g_state = 1;
static double urand() {  // SplitMix64 -> [0, 1)
  unsigned long long z = (g_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
static double sym() { return 2.0 * urand() - 1.0; }

template <int NN>
static int forward_mode() {
  ElemConst k;
  int M;
  if (scanf("%lf %lf %lf %lf %lf %d %d", &k.dt, &k.G, &k.K, &k.cs0, &k.hg, &k.hourglass, &M) != 7) return 2;
  // the law: 0 (bilinear), 1 Kh mh eps0 eps1 sy0m (Hollomon),
  // 2 B n C eps0 soft cap E (Johnson-Cook, A the per-element sy0) or
  // 3 ga gexp gn gm gi e_min e_max er_min er_max cap (GMT)
  if (scanf("%d", &k.law) != 1) return 2;
  if (k.law == 1 && scanf("%lf %lf %lf %lf %lf", &k.Kh, &k.mh, &k.eps0, &k.eps1, &k.sy0m) != 5) return 2;
  if (k.law == 2 && scanf("%lf %lf %lf %lf %lf %lf %lf", &k.jB, &k.jn, &k.jC, &k.jeps0, &k.jsoft, &k.jcap,
                          &k.Emod) != 7)
    return 2;
  if (k.law == 3 && scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &k.ga, &k.gexp, &k.gn, &k.gm, &k.gi,
                          &k.ge_min, &k.ge_max, &k.ger_min, &k.ger_max, &k.jcap) != 10)
    return 2;
  for (int e = 0; e < M; ++e) {
    ElemIn<NN> in;
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) scanf("%lf", &in.X[a][c]);
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) scanf("%lf", &in.V[a][c]);
    for (int a = 0; a < NN; ++a) scanf("%lf", &in.r[a]);
    for (int q = 0; q < 6; ++q) scanf("%lf", &in.tau[q]);
    if (scanf("%lf %lf %lf %lf", &in.ep, &in.rho, &in.sy0, &in.H) != 4) return 2;
    ElemFwd<NN> F;
    elem_forward<NN>(k, in, F);
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) printf("%.17g ", F.fnet[a][c]);
    for (int q = 0; q < 6; ++q) printf("%.17g ", F.t2[q]);
    printf("%.17g %.17g %.17g %d\n", F.ep2, F.p, F.vol, F.plastic);
  }
  return 0;
}

// scalar functional of the element outputs for the difference check
template <int NN>
static double phi(const ElemConst &k, const ElemIn<NN> &in, const double fb[NN][3], const double tb[6],
                  double epb, int *plastic) {
  ElemFwd<NN> F;
  elem_forward<NN>(k, in, F);
  double s = epb * F.ep2;
  for (int a = 0; a < NN; ++a)
    for (int c = 0; c < 3; ++c) s += fb[a][c] * F.fnet[a][c];
  for (int q = 0; q < 6; ++q) s += tb[q] * F.t2[q];
  if (plastic) *plastic = F.plastic;
  return s;
}

template <int NN>
static int vjp_mode(int count, int want_plastic, int law) {
  double worst = 0.0;
  int used = 0;
  for (int t = 0; t < count * 20 && used < count; ++t) {
    ElemConst k{1.0e-6, 76.9e9, 166.7e9, 4600.0, 0.06, NN == 8};
    if (law == 1) {  // power branch throughout (eps1 below every s)
      k.law = 1;
      k.mh = 0.1 + 0.3 * urand();
      k.eps0 = 1.5e-3;
      k.eps1 = 1.0e-4;
      k.sy0m = 2.0e8;
    } else if (law == 2) {  // Johnson-Cook, strain rates well below the cap
      k.law = 2;
      k.jn = 0.2 + 0.5 * urand();
      k.jC = 0.01 + 0.05 * urand();
      k.jeps0 = 1.0;
      k.jsoft = 0.7 + 0.2 * urand();
      k.jcap = 1.0e6;
      k.Emod = 200.0e9;
    } else if (law == 3) {  // GMT at 300 (the constants of the GPU material tests), ranges
      // wide enough that the clamps stay open
      k.law = 3;
      const double Tc = 300.0;
      k.gexp = exp(-1.5e-3 * Tc);
      k.gn = -2e-4 * Tc + 0.15 + 0.1 * urand();
      k.gm = 1e-4 * Tc + 0.01 + 0.02 * urand();
      k.gi = -1e-5 * Tc + 1e-4;
      k.ge_min = 1e-4;
      k.ge_max = 3.0;
      k.ger_min = 1e-3;
      k.ger_max = 1e5;
      k.jcap = 1.0e6;
    }
    ElemIn<NN> in;
    const double h = 0.02;
    for (int a = 0; a < NN; ++a) {
      const double base[3] = {NN == 8 ? (double)((((a + 1) >> 1) & 1)) : (a == 1 ? 1.0 : 0.0),
                              NN == 8 ? (double)((a >> 1) & 1) : (a == 2 ? 1.0 : 0.0),
                              NN == 8 ? (double)((a >> 2) & 1) : (a == 3 ? 1.0 : 0.0)};
      for (int c = 0; c < 3; ++c) {
        in.X[a][c] = h * (base[c] + 0.12 * sym());
        in.V[a][c] = 3.0 * sym();
      }
      in.r[a] = 1.0 + 2.0e-3 * sym();
    }
    for (int q = 0; q < 6; ++q) in.tau[q] = 2.0e8 * sym();
    // Johnson-Cook: ep^(n-2) curvature kept moderate away from ep = 0
    in.ep = law >= 2 ? 0.005 + 0.02 * urand() : 0.02 * urand();
    in.rho = 7850.0;
    in.H = 1.0e9 * (0.5 + urand());
    // yield stress well below or above the trial stress
    {
      ElemIn<NN> probe = in;
      probe.sy0 = 1.0e30;
      ElemFwd<NN> F;
      elem_forward<NN>(k, probe, F);
      in.sy0 = (want_plastic ? 0.6 : 1.6) * F.st - in.H * in.ep;
      if (law == 3) {  // C1 from the target yield stress
        const double target = (want_plastic ? 0.6 : 1.6) * F.st;
        const double es = in.ep, er = F.edot;
        k.ga = target / (pow(es, k.gn) * exp(k.gi / es) * pow(er, k.gm));
      }
      if (law == 2) {  // A and B ep^n in equal parts of the target yield stress
        const double lt = 1.0 + k.jC * log(F.edot / k.jeps0);
        const double target = (want_plastic ? 0.6 : 1.6) * F.st / (lt * k.jsoft);
        in.sy0 = 0.5 * target;
        k.jB = 0.5 * target / pow(in.ep, k.jn);
      }
      if (in.sy0 <= 0.0) continue;
      // Hollomon: Kh with Kh s^mh the same fraction of the trial stress
      if (law == 1) k.Kh = (want_plastic ? 0.6 : 1.6) * F.st / pow(in.ep + k.eps0, k.mh);
    }
    double fb[NN][3], tb[6], epb = 1.0e3 * sym();
    for (int a = 0; a < NN; ++a)
      for (int c = 0; c < 3; ++c) fb[a][c] = sym();
    for (int q = 0; q < 6; ++q) tb[q] = 1.0e-8 * sym();
    ElemFwd<NN> F;
    elem_forward<NN>(k, in, F);
    if (F.plastic != want_plastic) continue;
    ElemBar<NN> B;
    elem_reverse<NN>(k, in, F, fb, tb, epb, B);
    double Xv[NN][3];
    vol_reverse<NN>(in.X, B.volb, Xv);
    // directions: all inputs at once, then every parameter alone, then the
    // step length
    for (int dir = 0; dir < 7; ++dir) {
      ElemIn<NN> d{};
      double dhg = 0.0, dK = 0.0, dm = 0.0, dB = 0.0, dC1 = 0.0, dn = 0.0, ddt = 0.0;
      if (dir < 3) {
        for (int a = 0; a < NN; ++a) {
          for (int c = 0; c < 3; ++c) {
            d.X[a][c] = (dir == 0 || dir == 2) ? 0.05 * h * sym() : 0.0;
            d.V[a][c] = (dir == 1 || dir == 2) ? sym() : 0.0;
          }
          d.r[a] = (dir == 2) ? 1.0e-3 * sym() : 0.0;
        }
        for (int q = 0; q < 6; ++q) d.tau[q] = (dir == 2) ? 1.0e8 * sym() : 0.0;
        d.ep = (dir == 2) ? 0.01 * sym() : 0.0;
      } else if (dir == 3) {
        if (law == 1) dK = 1.0e-3 * k.Kh;
        else if (law == 3) dC1 = 1.0e-3 * k.ga / k.gexp;
        else d.sy0 = 1.0e8;
      } else if (dir == 4) {
        if (law == 1) dm = 0.01;
        else if (law == 2) dB = 1.0e-3 * k.jB;
        else if (law == 3) dn = 0.01;
        else d.H = 1.0e9;
      } else if (dir == 5) {
        dhg = 0.05;
      } else {
        ddt = k.dt;
      }
      // Johnson-Cook: the B row sits in B.H
      // GMT: the C1 row sits in B.sy0, the n2 row in B.H
      double ad = B.ep * d.ep + B.sy0 * (law == 3 ? dC1 : d.sy0) + B.H * (law == 2 ? dB : (law == 3 ? dn : d.H)) +
                  B.Kh * dK + B.mh * dm + B.hg * dhg + B.dt * ddt;
      for (int a = 0; a < NN; ++a) {
        for (int c = 0; c < 3; ++c) ad += (B.X[a][c] + Xv[a][c]) * d.X[a][c] + B.V[a][c] * d.V[a][c];
        ad += B.r[a] * d.r[a];
      }
      for (int q = 0; q < 6; ++q) ad += B.tau[q] * d.tau[q];
      // central differences: truncation ~eps^2 and rounding ~1/eps balance
      // near 1e-4 for these input scales (error 2e-8 there); the Hollomon
      // law's s^(mh-2) curvature at small plastic strain (s = ep + 1.5e-3)
      // needs the smaller step 3e-5 for the same balance
      const double eps = law == 1 ? 3.0e-5 : 1.0e-4;
      auto shifted = [&](double s, int *pl) {
        ElemIn<NN> p = in;
        ElemConst kk = k;
        for (int a = 0; a < NN; ++a) {
          for (int c = 0; c < 3; ++c) {
            p.X[a][c] += s * d.X[a][c];
            p.V[a][c] += s * d.V[a][c];
          }
          p.r[a] += s * d.r[a];
        }
        for (int q = 0; q < 6; ++q) p.tau[q] += s * d.tau[q];
        p.ep += s * d.ep;
        p.sy0 += s * d.sy0;
        p.H += s * d.H;
        kk.hg += s * dhg;
        kk.Kh += s * dK;
        kk.mh += s * dm;
        kk.jB += s * dB;
        kk.ga += s * dC1 * kk.gexp;  // ga = C1 exp(C2 Tc)
        kk.gn += s * dn;             // gn = n1 Tc + n2
        kk.dt += s * ddt;
        return phi<NN>(kk, p, fb, tb, epb, pl);
      };
      int p1 = 0, p2 = 0;
      const double fd = (shifted(eps, &p1) - shifted(-eps, &p2)) / (2.0 * eps);
      if (p1 != want_plastic || p2 != want_plastic) continue;  // crossed the yield surface
      const double scale = fmax(fabs(fd), fabs(ad));
      const double err = scale > 0 ? fabs(fd - ad) / scale : 0.0;
      if (scale > 0 && err > worst) worst = err;
      if (dir == 0) ++used;
    }
  }
  printf("%d %.3e\n", used, worst);
  return used == count ? 0 : 3;
}

int main(int argc, char **argv) {
  if (argc >= 3 && !strcmp(argv[1], "forward")) {
    int nn = atoi(argv[2]);
    return nn == 8 ? forward_mode<8>() : forward_mode<4>();
  }
  if (argc >= 6 && !strcmp(argv[1], "vjp")) {
    int nn = atoi(argv[2]);
    g_state = strtoull(argv[3], nullptr, 10);
    int count = atoi(argv[4]), pl = atoi(argv[5]), law = argc >= 7 ? atoi(argv[6]) : 0;
    return nn == 8 ? vjp_mode<8>(count, pl, law) : vjp_mode<4>(count, pl, law);
  }
  fprintf(stderr, "usage: forward NN < input | vjp NN SEED COUNT PLASTIC [LAW]\n");
  return 1;
}
