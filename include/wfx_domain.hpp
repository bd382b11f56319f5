// wfx_domain.hpp — header-only C++ mirror of MetFEM::Domain_d over the wfx
// C-ABI (include/wfx.h).  Method names, argument meaning and call order are
// the reference's (include/common/Domain_d.h:238-830), so the body of
// Domain_d::SolveChungHulbert (src/explicit/Solver_explicit.C:346-1170) can
// call these methods unchanged.  Every call runs HIP kernels on the stream
// given at construction; a failing call throws wfx::Error with the library's
// message.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "wfx.h"

namespace wfx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != WFX_OK) throw Error(rc, std::string("wfx: ") + wfx_last_error());
}

struct double3 {
  double x, y, z;
};
inline double3 make_double3(double x, double y, double z) { return {x, y, z}; }

class Domain_d {
 public:
  explicit Domain_d(int precision = WFX_FP64, wfx_stream stream = nullptr) : s_(stream) {
    check(wfx_create(&c_, precision));
  }
  ~Domain_d() { wfx_destroy(c_); }
  Domain_d(const Domain_d &) = delete;
  Domain_d &operator=(const Domain_d &) = delete;

  wfx_ctx *handle() { return c_; }

  // ---- mesh (Domain_d.C:1136-1504, :1508-1699) ----
  void AddBoxLength(const double3 &V, const double3 &L, const double &r, const bool &red_int = true,
                    const bool &tritetra = false) {
    const double v[3] = {V.x, V.y, V.z}, l[3] = {L.x, L.y, L.z};
    check(wfx_add_box_length(c_, v, l, r, red_int, tritetra));
  }
  void CreateFromMesh(const std::vector<double> &x, const std::vector<int> &elnod, int nodxelem) {
    check(wfx_set_mesh(c_, (int)(x.size() / 3), (int)(elnod.size() / nodxelem), nodxelem, x.data(),
                       elnod.data()));
  }
  int getNodeCount() { int n, e, k; check(wfx_get_sizes(c_, &n, &e, &k)); return n; }
  int getElemCount() { int n, e, k; check(wfx_get_sizes(c_, &n, &e, &k)); return e; }

  // ---- set-up ----
  void setDensity(const double &r) { check(wfx_set_density(c_, r)); }
  // Material_ + AssignMaterial: model WFX_BILINEAR (p1 = Et) or WFX_HOLLOMON (p1 = K, p2 = m)
  void AssignMaterial(int model, double E, double nu, double rho0, double sy0, double p1 = 0.0,
                      double p2 = 0.0) {
    check(wfx_set_material(c_, model, E, nu, rho0, sy0, p1, p2));
  }
  // JohnsonCook(el, A, B, n, C, eps_0, m, T_m, T_t) + AssignMaterial
  void AssignJohnsonCook(double E, double nu, double rho0, double A, double B, double n, double C,
                         double eps_0, double m, double T_m, double T_t) {
    check(wfx_set_material_jc(c_, E, nu, rho0, A, B, n, C, eps_0, m, T_m, T_t));
  }
  // GMT(el, n1, n2, C1, C2, m1, m2, I1, I2, ranges) + AssignMaterial
  void AssignGMT(double E, double nu, double rho0, double sy0, const double c[8], const double range[6]) {
    check(wfx_set_material_gmt(c_, E, nu, rho0, sy0, c, range));
  }
  void setTemp(const double &T) { check(wfx_set_temperature(c_, T)); }
  void setElemTemp(const std::vector<double> &T) {
    check(wfx_set_elem_temperature(c_, T.data(), (int)T.size()));
  }
  void setMaxStrainRate(const double &edot) { check(wfx_set_max_strain_rate(c_, edot)); }
  // setTriMesh + setContactOn: rigid triangle surfaces for penalty contact
  void setTriMesh(const std::vector<double> &nodes, const std::vector<int> &tris,
                  const std::vector<double> &node_vel = {}, const std::vector<int> &mesh_id = {}) {
    check(wfx_set_rigid_mesh(c_, (int)(nodes.size() / 3), nodes.data(), node_vel.empty() ? nullptr : node_vel.data(),
                             (int)(tris.size() / 3), tris.data(), mesh_id.empty() ? nullptr : mesh_id.data()));
    contact_ = true;
  }
  void setContactPF(double pf, double mu_static = 0.0, double mu_dynamic = 0.0) {
    check(wfx_set_contact_params(c_, pf, mu_static, mu_dynamic));
  }
  // heatCondCoeff and dieTemp of the contact block (main.C:718-719)
  void setContactHeat(double heat_cond, double die_temp) { check(wfx_set_contact_heat(c_, heat_cond, die_temp)); }
  double contactForce(int mesh_id = 0) { double f; check(wfx_contact_force(c_, mesh_id, &f)); return f; }
  void CalcExtFaceAreas() { check(wfx_calc_ext_face_areas(c_, s_)); }
  void CalcContactForces() { check(wfx_calc_contact_forces(c_, s_)); }
  // setThermalOn with the material's thermal constants
  void setThermal(bool on, double k_T, double cp_T, double plheat_fraction = 0.9) {
    check(wfx_set_thermal(c_, on ? 1 : 0, k_T, cp_T, plheat_fraction));
    thermal_ = on;
  }
  void ThermalCalcs() { check(wfx_thermal_calcs(c_, s_)); }
  void AddBCVelNode(const int &node, const int &dim, const double &val) {
    check(wfx_add_bc_vel_node(c_, node, dim, val));
  }
  void AllocateBCs() { check(wfx_allocate_bcs(c_)); }
  void SetDT(const double &dt) { dt_ = dt; check(wfx_set_dt(c_, dt)); }
  void setCFL(const double &cfl) { cfl_ = cfl; check(wfx_set_cfl(c_, cfl)); }
  // the per-step CFL time step of the current state (Solver_explicit.C:576-600)
  double calcCflDt() { check(wfx_calc_cfl_dt(c_, &dt_, s_)); return dt_; }
  void SetEndTime(const double &t) { end_t_ = t; }
  void setParam(const char *name, double v) { check(wfx_set_param(c_, name, v)); }
  void setReactionNodes(const std::vector<int> &nodes) {
    check(wfx_set_reaction_nodes(c_, nodes.data(), (int)nodes.size()));
  }
  double calcMinEdgeLength() { double v; check(wfx_calc_min_edge_length(c_, &v, s_)); return v; }
  double getDT() const { return dt_; }

  // ---- init block and per-step API (Solver_explicit.C order) ----
  void InitSolve() { check(wfx_init(c_, s_)); step_ = 0; }
  // The pieces of SolveChungHulbert's init block (Solver_explicit.C:101-340):
  // InitValues (Domain_d.C:393-416) runs the whole block (wfx_init: zero
  // state, BC velocities, Jacobians, vol0 = vol, rho = rho0, nodal volume and
  // mass, sigma_y = sy0); the other three are part of it and do nothing here.
  void InitValues() { InitSolve(); }
  void AssignMatAddress() {}      // Domain_d.C:945-949
  void CalcElemInitialVol() {}    // Mechanical.C:343-357
  void calcElemDensity() {}       // Mechanical.C:295-319
  // computeEnergies (Mechanical.C:2145-2185): kinetic energy and the internal
  // energy of the step, dEint = dt sum sigma:D vol; the reference leaves Evisc
  // unassigned, here it is 0
  void computeEnergies(double dt, double &Ekin, double &dEint, double &Evisc) {
    check(wfx_energies(c_, &Ekin, &dEint, s_));
    if (dt != dt_ && dt_ > 0.0) dEint = dEint / dt_ * dt;  // wfx_energies uses the domain's dt
    Evisc = 0.0;
  }
  // SearchExtNodes (Domain_d.C:110-296): the external faces are found when the
  // rigid surface is set; this refreshes the nodal areas
  void SearchExtNodes() { CalcExtFaceAreas(); }
  void UpdatePrediction() { check(wfx_update_prediction(c_, s_)); }
  void ImposeBCV(const int dim) { check(wfx_impose_bcv(c_, dim, s_)); }
  void ImposeBCA(const int dim) { check(wfx_impose_bca(c_, dim, s_)); }
  void ImposeBCVAllDim() { for (int d = 0; d < 3; ++d) ImposeBCV(d); }
  void ImposeBCAAllDim() { for (int d = 0; d < 3; ++d) ImposeBCA(d); }
  void calcElemJAndDerivatives() { check(wfx_calc_elem_jacobians(c_, s_)); }
  void CalcElemVol() { check(wfx_calc_elem_vol(c_, s_)); }
  void CalcNodalVol() { check(wfx_calc_nodal_vol(c_, s_)); }
  void CalcNodalMassFromVol() { check(wfx_calc_nodal_mass_from_vol(c_, s_)); }
  void calcElemStrainRates() { check(wfx_calc_elem_strain_rates(c_, s_)); }
  void calcElemPressure() { check(wfx_calc_elem_pressure(c_, s_)); }
  void CalcStressStrain(double dt) { check(wfx_calc_stress_strain(c_, dt, s_)); }
  void calcArtificialViscosity() { check(wfx_calc_artificial_viscosity(c_, s_)); }
  void calcElemForces() { check(wfx_calc_elem_forces(c_, s_)); }
  void calcElemHourglassForces() { check(wfx_calc_elem_hourglass_forces(c_, s_)); }
  void assemblyForces() { check(wfx_assembly_forces(c_, s_)); }
  void calcAccel() { check(wfx_calc_accel(c_, s_)); }
  void UpdateCorrectionAccVel() { check(wfx_update_correction_accvel(c_, s_)); }
  void UpdateCorrectionPos() { check(wfx_update_correction_pos(c_, s_)); }
  // sum over the first reaction group; reactionForces gives every group
  double reactionForce(int dim = 2) { return reactionForces(dim)[0]; }
  std::vector<double> reactionForces(int dim = 2) {
    int ng = 0;
    check(wfx_reaction_group_count(c_, &ng));
    std::vector<double> r(ng > 0 ? ng : 1);
    check(wfx_reaction_force(c_, dim, r.data(), s_));
    return r;
  }

  // ---- specimen batches, launch tuning, graphs (no reference counterpart) ----
  void setReactionGroups(const std::vector<int> &nodes, const std::vector<int> &offsets) {
    check(wfx_set_reaction_groups(c_, nodes.data(), (int)nodes.size(), offsets.data(),
                                  (int)offsets.size() - 1));
  }
  void setElemTangent(const std::vector<double> &Et) {
    check(wfx_set_elem_tangent(c_, Et.data(), (int)Et.size()));
  }
  std::vector<int> autotune(int reps = 5) {
    std::vector<int> bs(4);
    check(wfx_autotune(c_, reps, bs.data(), s_));
    return bs;
  }
  void graphCapture(int nsteps, bool reaction = true) { check(wfx_graph_capture(c_, nsteps, reaction, s_)); }
  void graphLaunch(double *reaction_hist = nullptr) { check(wfx_graph_launch(c_, reaction_hist, s_)); }

  // The CPU-branch loop body of SolveChungHulbert written with the per-step
  // calls (remeshing, contact, thermal and output left out).
  void StepChungHulbert() {
    if (contact_ && step_ % 10 == 0) CalcExtFaceAreas();
    UpdatePrediction();
    ImposeBCVAllDim();
    calcElemJAndDerivatives();
    if (cfl_ > 0.0) calcCflDt();
    CalcElemVol();
    CalcNodalVol();
    CalcNodalMassFromVol();
    calcElemStrainRates();
    calcElemPressure();
    CalcStressStrain(dt_);
    calcArtificialViscosity();
    calcElemForces();
    calcElemHourglassForces();
    if (contact_) CalcContactForces();
    assemblyForces();
    calcAccel();
    ImposeBCAAllDim();
    UpdateCorrectionAccVel();
    ImposeBCVAllDim();
    UpdateCorrectionPos();
    if (contact_) check(wfx_rigid_update(c_, s_));
    if (thermal_) ThermalCalcs();
    ++step_;
  }

  // SolveChungHulbert without output: init, then loop while Time < end_t.
  // fused = true runs the same sequence through wfx_step_fused.
  int SolveChungHulbert(bool fused = false) {
    InitSolve();
    double Time = 0.0;
    int steps = 0;
    while (Time < end_t_) {
      if (fused) {
        check(wfx_step_fused(c_, 1, nullptr, s_));
        check(wfx_get_dt(c_, &dt_));
      } else {
        StepChungHulbert();
      }
      Time += dt_;
      ++steps;
    }
    return steps;
  }

  std::vector<double> field(const char *name) {
    long n = 0;
    int is_int = 0;
    check(wfx_field_count(c_, name, &n, &is_int));
    if (is_int) throw Error(WFX_ERR_ARG, std::string("integer field ") + name);
    std::vector<double> out(n);
    check(wfx_get_field(c_, name, out.data(), n));
    return out;
  }

 private:
  wfx_ctx *c_ = nullptr;
  wfx_stream s_ = nullptr;
  double dt_ = 0.0, end_t_ = 0.0, cfl_ = 0.0;
  bool thermal_ = false, contact_ = false;
  long step_ = 0;
};

}  // namespace wfx
