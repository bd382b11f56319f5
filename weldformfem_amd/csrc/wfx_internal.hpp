// Internal definitions shared by the wfx host code and the HIP kernels.
//
// Device data layout (HBM), structure of arrays throughout:
//   nodal vectors      double  [3][N]      component-major (x_x..., x_y..., x_z...)
//   nodal scalars      double  [N]
//   element scalars    real    [E]         real = double (FP64) or float (FP32-mixed)
//   element tensors    real    [6][E]      components xx,yy,zz,xy,yz,xz
//   element spin       real    [3][E]      components xy,yz,xz (diagonal is zero)
//   shape derivatives  real    [3][n][E]   dH_detJ(d,a) at row d*n+a
//   element forces     real    [n][3][E]   f(a,d) at row a*3+d
//   connectivity       int32   [n][E]      elnod(e,a) at a*E+e
//   node->element CSR  int32   nodel[nnz], nodel_loc[nnz], offset[N], count[N]
// Element-thread kernels therefore read and write every element field with
// unit stride across a wavefront; nodal gathers go through L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#define WFX_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      wfx::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));      \
      return WFX_ERR_HIP;                                                     \
    }                                                                         \
  } while (0)

namespace wfx {

void set_error(const std::string &msg);
struct Adjoint;  // checkpointed adjoint run (wfx_adjoint.hip)
struct Frames;   // asynchronous VTK frames (wfx_frames.hip)

enum Model { BILINEAR = 0, HOLLOMON = 1, JOHNSON_COOK = 2, GMT = 3 };

// slots of the per-step CFL time step in the context's block of 64 device
// scalars: minimum edge length, maximum speed, sound speed sqrt(K / rho[0])
// and dt of the last step
constexpr int kCflMinl = 16, kCflVmax = 17, kCflCs = 18, kCflDt = 20;

struct Material {
  int model = BILINEAR;
  double E = 0, nu = 0, K = 0, G = 0, rho0 = 0, cs0 = 0, sy0 = 1.0e10;
  double Et = 0, H = 0;                 // bilinear: tangent, plastic modulus
  double Kh = 0, mh = 1, eps0 = 0, eps1 = 0;  // Hollomon
  double B = 0, n = 1, C = 0, eps_0 = 1, m = 1, T_m = 1e10, T_t = 0;  // Johnson-Cook (A = sy0)
  double C1 = 0, C2 = 0, n1 = 0, n2 = 0, m1 = 0, m2 = 0, I1 = 0, I2 = 0;  // GMT
  double e_min = 0, e_max = 1e10, er_min = 0, er_max = 1e10, T_min = 0, T_max = 1e10;
  double temp = 20.0;      // uniform element temperature (IniTemp, main.C:424)
  double max_edot = 1e6;   // strain-rate cap of the tangent (Domain_d.h:824)
};

struct Stab {
  double alpha_free = 0, log_factor = 0, J_min = 0, hg_coeff_free = 0;
  double pspg_scale = 0, p_pspg_bulkfac = 0, av_coeff_div = 0, av_coeff_bulk = 0;
  double alpha_contact = 0, hg_coeff_contact = 0;  // elements with a node in contact
  double artifvisc0 = 0, artifvisc1 = 0;
  double hg_coeff = 0;  // 3D hexa viscous hourglass coefficient
};

// elements per cluster of the fused element pass (its work-group size)
constexpr int kCluster = 256;

// slots of the context's device status counters (unsigned int[16])
constexpr int kStatNonfinite = 0;  // non-finite internal-force entries scrubbed
constexpr int kStatInverted = 1;   // element Jacobians with detJ <= 0 or not finite

// Device buffers of one domain.  T is the element-field real type.
template <typename T>
struct DevElem {
  T *detJ = nullptr, *dH = nullptr, *vol = nullptr, *vol0 = nullptr, *rho = nullptr,
    *rho0 = nullptr, *p = nullptr, *pl_strain = nullptr, *sigma_y = nullptr, *sy0 = nullptr;
  T *str_rate = nullptr, *rot_rate = nullptr, *sigma = nullptr, *tau = nullptr, *eps = nullptr,
    *strain_pl_incr = nullptr;
  T *f_elem = nullptr, *f_hg = nullptr;
  T *He = nullptr;  // optional per-element bilinear plastic modulus (specimen batches)
  T *q_plheat = nullptr, *dTedt = nullptr;  // thermal coupling: plastic heat, element-node rates [n][E]
  T *avq = nullptr;  // fused path: the artificial-viscosity term added to the diagonal stresses
  T *part = nullptr;  // fused path: per-slot partial nodal forces [4][S] (fx, fy, fz, thermal term)
  // thermal expansion: the element-node temperature rates of the last step at
  // the reference's flat indices [e*n + a] below N (the entries
  // calcThermalExpansion reads, at node ids), and the fused path's
  // per-element mean of them [E]
  T *dte_flat = nullptr, *texp = nullptr;
};

// Rigid-surface penalty contact (Contact.C, TriMesh_d): penalty factor
// (m_contPF), friction coefficients, sizes, the solver's step counter (the
// external face areas are refreshed every 10 steps)
struct Contact {
  int on = 0;
  double pf = 0.1, mu_sta = 0.0, mu_dyn = 0.0;
  int rn = 0, rt = 0, next = 0;
  long step_count = 0;
  // device
  double *rnode = nullptr, *rvel = nullptr, *rnormal = nullptr, *rpplane = nullptr;  // [n][3] / [t][3] / [t]
  int *rtri = nullptr, *rmesh = nullptr;
  int *ext_nodes = nullptr;                      // external node ids
  int *afc_off = nullptr, *afc_face = nullptr;   // external node k: its external faces, in face order
  int *ext_face = nullptr;                       // [nf][3], node order of the element's face table
  double *node_area = nullptr, *elem_length = nullptr;
  double *contforce = nullptr, *ut_prev = nullptr;  // [3][N], kept between steps
  // contact heat conduction (heatCondCoeff, dieTemp: main.C:714-719): the
  // conduction into each node in contact, kept between steps [N]
  double heat_cond = 0.0, T_const = 20.0;
  double *q_cont = nullptr;
  int *mesh_in_contact = nullptr;
  double *cfn = nullptr;
};

// Thermal coupling (setThermalOn, Thermal.C): conductivity, heat capacity,
// fraction of plastic work turned into heat (plHeatFrac)
struct Thermal {
  int on = 0;
  double k_T = 0, cp_T = 0, plheat_frac = 0.9;
  double exp_T = 0;  // thermal expansion coefficient (Material_::exp_T); 0: none
};

struct DevNode {
  double *x = nullptr, *v = nullptr, *a = nullptr, *u = nullptr, *u_dt = nullptr,
         *prev_a = nullptr, *fi = nullptr, *fe = nullptr;
  double *T = nullptr;  // nodal temperature (thermal coupling)
  double *mdiag = nullptr, *voln = nullptr, *voln0_sum = nullptr;
  double *anp_ratio = nullptr;  // fused path: V_n / V0_n
};

struct DevMesh {
  int *elnod = nullptr;
  int *nodel = nullptr, *nodel_loc = nullptr, *nodel_offset = nullptr, *nodel_count = nullptr;
  int *nodel_pk = nullptr;  // (element << 3) | local node, CSR order
  int *nodel_pad = nullptr; // the first 8 packed entries of every node, [8][N], -1 past its count
  // element clusters of the fused element pass (wfx_cluster.hpp): cluster
  // element and slot offsets [C+1], per-cluster slot entry offsets [S+C],
  // element-node entries [E*n], node -> slots CSR offsets [N+1]; FP64: the
  // slots of that CSR [S]; FP32-mixed: each slot's position in it [S] (the
  // partial sums are stored there, node-major)
  int ncl = 0;
  long nslot = 0;
  int *cl_eoff = nullptr, *cl_sbase = nullptr;
  uint16_t *cl_loff = nullptr, *cl_ent = nullptr;
  int *nsl_off = nullptr, *nsl = nullptr, *spos = nullptr;
  // slot -> node [S]; each stored element's nodes as slot indices of its cluster [E][n]
  int *cl_snode = nullptr;
  uint16_t *cl_lslot = nullptr;
  int *flat_base = nullptr;  // thermal expansion: original element * n of each stored element
};

// Kernel launch parameters shared by the precision variants.
struct StepParams {
  int N = 0, E = 0, n = 0;
  double dt = 0, alpha = 0, beta = 0, gamma = 0;
  double cfl = 0;  // > 0: per-step CFL time step with this factor (dt is the initial one)
  Material mat;
  Stab stab;
};

}  // namespace wfx

// The opaque context behind the C-ABI handle.
struct wfx_ctx {
  int precision = 0;  // 0 FP64, 1 FP32-mixed
  int N = 0, E = 0, n = 0, nnz = 0;
  bool has_mesh = false, initialized = false;
  wfx::StepParams prm;
  // host copies of the mesh and boundary conditions
  std::vector<double> x_h;       // reference layout [N][3]
  std::vector<int> elnod_h;      // [E][n]
  // element storage order on the device (cluster order): stored index s holds
  // original element el_s2o[s]; every element field crosses the boundary in
  // the original order
  std::vector<int> el_s2o, el_o2s;
  std::vector<int> nodel_h, nodel_loc_h, nodel_offset_h, nodel_count_h;
  std::vector<double> sy0_h;     // per element
  std::vector<int> bc_nod_h[3];
  std::vector<double> bc_val_h[3];
  std::vector<int> react_h;
  // device
  wfx::DevNode nd;
  wfx::DevMesh ms;
  wfx::DevElem<double> e64;
  wfx::DevElem<float> e32;
  int *bc_nod[3] = {nullptr, nullptr, nullptr};
  double *bc_val[3] = {nullptr, nullptr, nullptr};
  int bc_count[3] = {0, 0, 0};
  int *react = nullptr;             // reaction nodes, grouped
  int *react_off = nullptr;         // group g = react[react_off[g] .. react_off[g+1])
  std::vector<int> react_off_h;
  int ngroups = 0;
  // per-node Dirichlet data for the fused node pass: slot or -1, bit mask of
  // constrained directions and the three prescribed values per slot
  int *bc_slot = nullptr, *bc_mask = nullptr;
  double *bc_tab = nullptr;
  // FP32-mixed fused path: narrowed nodal state during a run, start positions
  float *v32 = nullptr, *pa32 = nullptr, *ud32 = nullptr, *fi32 = nullptr;
  double *x_start = nullptr;  // (both precisions)
  float *x_lo = nullptr;      // FP64 fused runs: rounding error of the position updates
  int has_fe = 0;            // external nodal forces set (non-zero)
  double rho_uniform = 0.0;  // > 0: every element density equals this value
  // host scan of vol0 and rho (wfx_init, wfx_set_field): elem_guard 1 when an
  // initial volume is not finite or (nearly) zero or a density is not positive
  // and finite, or when they changed on the device since the last scan;
  // rho_min is the smallest density found
  int elem_guard = 1;
  double rho_min = 0.0;
  wfx::Thermal thermal;
  wfx::Contact contact;
  int graph_phase = 0;       // step counter mod 10 when the graph was captured
  double *temp_e = nullptr;  // per-element temperature (else prm.mat.temp for all)
  double *temp_buf = nullptr;  // device buffer behind temp_e, kept once allocated
  int fused_bs[5] = {256, 256, 256, 256, 256};  // block size of fused phases 1..4
  // captured fused-step graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 0;
  bool graph_hist = false;
  uint64_t graph_sig = 0;
  bool fused_stale = false;  // per-step intermediates not maintained by the fused path
  int fused_write_fi = 1;    // the fused node pass stores the internal force (cleared inside runs)
  int fused_write_out = 1;   // the fused element pass stores p, sigma_y, plastic heat (cleared inside runs)
  int fused_read_ud = 1;     // the fused node pass reads the predicted increment u_dt (cleared inside runs)
  int fused_write_ud = 1;    // and stores it every step (cleared inside runs of a fixed time step)
  int fused_run_first = 1;   // the fused step is the first of a run
  int fused_run_last = 1;    // the fused step is the last of a run
  int f2_lanes = 0;          // the nodal volume sums take several lanes per node (many elements per node)
  double *scalars = nullptr;      // device scratch for reductions (reaction force)
  // this context's device's address of the library's static scratch array
  // (the allocation-free entry point's reduction partials)
  double *scratch = nullptr;
  unsigned int *status = nullptr; // counters, slots kStat*
  unsigned int status_seen[2] = {0, 0};  // counters at the end of the last checked run
  int check_status = 1;           // runs read the counters at their end (synchronising)
  double *cfl_part = nullptr;     // per-block partial extremes of the CFL time step:
  int cfl_cap_min = 0;            //   edge minima at [0, cfl_cap_min), speed maxima after
  int cfl_nvmax = 0;              //   speed maxima left by the last node pass
  double *dt_hist = nullptr;      // per-step dt of the last run (per-step CFL time step)
  int dt_cap = 0;
  std::vector<double> dt_hist_h;  // host copy of the last run's per-step dt
  double *hist = nullptr;         // reaction-force history of wfx_run
  int hist_cap = 0;
  std::vector<void *> allocations;
  wfx::Adjoint *adj = nullptr;  // owned, freed with the context or a new mesh
  wfx::Frames *frames = nullptr;  // owned, joined and freed with the context or a new mesh
};
