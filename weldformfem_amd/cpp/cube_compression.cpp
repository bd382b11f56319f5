// C++ driver: the reference's C++-API recipe (src/common/main_1_elem_3d.C
// style) for a cube compression, solved through the Domain_d mirror.
//   wfx_cube <n> <steps> [fused]
// n^3 reduced-integration hexa in a 1 m cube, steel with bilinear hardening
// (E 200 GPa, nu 0.3, rho 7850, sy0 300 MPa, Et 1 GPa), bottom face fixed in
// z, top face at vz = -1 m/s, hourglass 0.06, dt = 0.6 * h / c_d.  Prints the
// reaction force history every 100 steps and the final top displacement.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "wfx_domain.hpp"

int main(int argc, char **argv) {
  int n = argc > 1 ? std::atoi(argv[1]) : 20;
  int steps = argc > 2 ? std::atoi(argv[2]) : 1000;
  bool fused = argc > 3 && std::atoi(argv[3]) != 0;
  try {
    wfx::Domain_d dom(WFX_FP64);
    const double L = 1.0, r = 0.5 * L / n;
    dom.AddBoxLength(wfx::make_double3(0, 0, 0), wfx::make_double3(L, L, L), r, true, false);
    const double E = 200e9, nu = 0.3, rho = 7850.0;
    dom.setDensity(rho);
    dom.AssignMaterial(WFX_BILINEAR, E, nu, rho, 300e6, 1e9);
    const int layer = (n + 1) * (n + 1);
    std::vector<int> top;
    for (int k = 0; k < layer; ++k) {
      dom.AddBCVelNode(k, 2, 0.0);
      dom.AddBCVelNode(n * layer + k, 2, -1.0);
      top.push_back(n * layer + k);
    }
    dom.AllocateBCs();
    dom.setReactionNodes(top);
    dom.setParam("hg_coeff", 0.06);
    const double K = E / (3 * (1 - 2 * nu)), G = E / (2 * (1 + nu));
    const double cd = std::sqrt((K + 4 * G / 3) / rho);
    const double dt = 0.8 * 0.75 * dom.calcMinEdgeLength() / cd;
    dom.SetDT(dt);
    // the init block under the reference's names (Solver_explicit.C:101-340)
    dom.AssignMatAddress();
    dom.InitValues();
    dom.CalcElemInitialVol();
    dom.calcElemDensity();
    for (int s = 0; s < steps; ++s) {
      if (fused) {
        wfx::check(wfx_step_fused(dom.handle(), 1, nullptr, nullptr));
      } else {
        dom.StepChungHulbert();
      }
      if ((s + 1) % 100 == 0 || s + 1 == steps) {
        std::printf("step %d reaction_z %.17g\n", s + 1, dom.reactionForce(2));
      }
    }
    if (!fused) {  // the per-step path keeps the strain rates energies need
      double ekin = 0, deint = 0, evisc = 0;
      dom.computeEnergies(dt, ekin, deint, evisc);
      std::printf("energies %.17g %.17g\n", ekin, deint);
    }
    std::vector<double> u = dom.field("u");
    std::vector<double> ep = dom.field("pl_strain");
    double umin = 0, epmax = 0;
    for (size_t i = 2; i < u.size(); i += 3) umin = std::fmin(umin, u[i]);
    for (double v : ep) epmax = std::fmax(epmax, v);
    std::printf("dt %.17g min_uz %.17g max_pl_strain %.17g\n", dt, umin, epmax);
  } catch (const wfx::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
