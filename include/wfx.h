/*
 * wfx.h — C-ABI of the MI355X-native explicit updated-Lagrangian step.
 *
 * Drop-in boundary for WeldFormFEM's per-step Domain_d API: every per-step
 * entry point below replaces one MetFEM::Domain_d member function (cited as
 * reference file:line) and runs as HIP kernels on the given stream.  Plain
 * pointers and sizes only; every call returns 0 on success or a negative
 * WFX_ERR_* code, with a message in wfx_last_error().
 *
 * State is owned by the context (device structure-of-arrays buffers).  Host
 * arrays passed to wfx_set_field / wfx_get_field use the reference's own
 * layout (Domain_d.h:841-1043): nodal vectors [N][3], element tensors
 * [E][6] as xx,yy,zz,xy,yz,xz, element forces [E][n][3], shape derivatives
 * [E][n] per direction.
 *
 * Threading: one context is used by one host thread at a time; kernels are
 * asynchronous on the stream passed in (NULL = default stream).
 */
#ifndef WFX_H
#define WFX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wfx_ctx wfx_ctx;
typedef void *wfx_stream; /* hipStream_t */

enum {
  WFX_OK = 0,
  WFX_ERR_ARG = -1,      /* bad argument or size                         */
  WFX_ERR_STATE = -2,    /* call out of order (no mesh, not initialised) */
  WFX_ERR_HIP = -3,      /* HIP runtime error                            */
  WFX_ERR_NONFINITE = -4 /* non-finite value detected                    */
};

enum { WFX_FP64 = 0, WFX_FP32_MIXED = 1 };
enum { WFX_BILINEAR = 0, WFX_HOLLOMON = 1, WFX_JOHNSON_COOK = 2, WFX_GMT = 3 };

/* Error message and library version: no reference counterpart (the
 * reference's methods return nothing and print their diagnostics). */
const char *wfx_last_error(void);
int wfx_version(void);

/* ---- lifetime ------------------------------------------------------------ */
/* new Domain_d (Domain_d.h:238); precision WFX_FP64 or WFX_FP32_MIXED */
int wfx_create(wfx_ctx **out, int precision);
/* Domain_d::Free (Domain_d.C:786-900) */
int wfx_destroy(wfx_ctx *ctx);

/* ---- mesh ---------------------------------------------------------------- */
/* Domain_d::AddBoxLength (Domain_d.C:1136-1504), 3D, reduced integration:
 * structured hexa (tritetra=0) or six tets per cube (tritetra=1). */
int wfx_add_box_length(wfx_ctx *ctx, const double V[3], const double L[3], double r,
                       int red_int, int tritetra);
/* Domain_d::CreateFromLSDyna + setNodElem (Domain_d.C:1647-1699, :1508-1611):
 * x[N][3], elnod[E][nodxelem] (nodxelem 4 or 8); builds the node->element CSR.
 * Both mesh entry points refuse (WFX_ERR_ARG) meshes of 2^28 elements or more
 * or of more than 2^31 - 1 element-node entries. */
int wfx_set_mesh(wfx_ctx *ctx, int node_count, int elem_count, int nodxelem,
                 const double *x, const int *elnod);
/* getNodeCount / getElemCount / m_nodxelem (Domain_d.h) */
int wfx_get_sizes(wfx_ctx *ctx, int *node_count, int *elem_count, int *nodxelem);

/* ---- set-up -------------------------------------------------------------- */
/* Domain_d::setDensity (Domain_d.C:952-958) */
int wfx_set_density(wfx_ctx *ctx, double rho);
/* Material_ + Elastic_ + AssignMaterial (Material.cuh:15-104, Domain_d.C:945).
 * model WFX_BILINEAR: p1 = tangent modulus Et (sigma_y = sy0 + H*ep with
 * H = E*Et/(E-Et)); model WFX_HOLLOMON: p1 = K, p2 = m (Material.cuh:353-395). */
int wfx_set_material(wfx_ctx *ctx, int model, double E, double nu, double rho0, double sy0,
                     double p1, double p2);
/* JohnsonCook(el, A, B, n, C, eps_0, m, T_m, T_t) + AssignMaterial
 * (Material.cuh:191-203, main.C:535-541; yield and tangent
 * Material.cuh:378-411): sigma_y = (A + B ep^n)(1 + C ln(edot/eps_0))
 * (1 - T_h^m), T_h = (T - T_t)/(T_m - T_t).  A is the per-element initial
 * yield (wfx_set_elem_sy0 overrides it element by element). */
int wfx_set_material_jc(wfx_ctx *ctx, double E, double nu, double rho0, double A, double B,
                        double n, double C, double eps_0, double m, double T_m, double T_t);
/* GMT(el, n1, n2, C1, C2, m1, m2, I1, I2, e_min, e_max, er_min, er_max, T_min,
 * T_max) + AssignMaterial (Material.cuh:271-291, main.C:548-553; yield and
 * tangent Material.cuh:412-473).  c = {n1, n2, C1, C2, m1, m2, I1, I2} in the
 * reference's input order, range = {e_min, e_max, er_min, er_max, T_min,
 * T_max}; sy0 is the initial sigma_y (yieldStress0). */
int wfx_set_material_gmt(wfx_ctx *ctx, double E, double nu, double rho0, double sy0,
                         const double c[8], const double range[6]);
/* Domain_d::setTemp (Domain_d.h:678-683, IniTemp main.C:424-440, default 20):
 * uniform temperature read by the Johnson-Cook and GMT laws, and the value of
 * every nodal temperature (the thermal coupling of wfx_set_thermal advances
 * those). */
int wfx_set_temperature(wfx_ctx *ctx, double T);
/* per-element temperatures (the reference's T[e] of CalcStressStrain,
 * Mechanical.C:1725-1729); NULL returns to the uniform value */
int wfx_set_elem_temperature(wfx_ctx *ctx, const double *T, int elem_count);
/* Domain_d::setThermalOn + the material's thermalCond (k_T), thermalHeatCap
 * (cp_T) and the domain's plHeatFrac (main.C:218, :590-592, :436-441):
 * thermal coupling of the BUILD_REMESH / CPU solver.  Each step then stores
 * the plastic heat source of CalcStressStrain (Mechanical.C:1784-1810) and
 * runs ThermalCalcs after the position update (Thermal.C:27-115,
 * Solver_explicit.C:1008-1012): element conduction on the step's Jacobians,
 * heat split evenly over the element's nodes, T += dTdt * dt per node.  The
 * Johnson-Cook and GMT laws then read the mean nodal temperature of each
 * element.  Contact conduction and thermal expansion are not included.
 * Nodal temperatures: field "T" (starting at wfx_set_temperature's value). */
int wfx_set_thermal(wfx_ctx *ctx, int on, double k_T, double cp_T, double plheat_fraction);
/* Domain_d::ThermalCalcs (Thermal.C:27-115) on the state of the last
 * per-step API step */
int wfx_thermal_calcs(wfx_ctx *ctx, wfx_stream stream);
/* Material_::exp_T, the material's "thermalExp" (main.C:577): with thermal
 * coupling on, every step then runs calcThermalExpansion after the strain
 * rates (Solver_explicit.C:719-720).  0 (the default) leaves the strain rate
 * unchanged. */
int wfx_set_thermal_expansion(wfx_ctx *ctx, double exp_T);
/* Domain_d::calcThermalExpansion (Thermal.C:164-178) on the strain rates of
 * the last per-step API step: D_ii -= exp_T * (mean over the element's nodes
 * of the element-node temperature rates at the flat index [e*n + a] equal to
 * the node id, which is how the reference indexes m_dTedt) */
int wfx_calc_thermal_expansion(wfx_ctx *ctx, wfx_stream stream);
/* Rigid-surface penalty contact (Domain_d::CalcContactForces,
 * Contact.C:31-335; TriMesh_d, Mesh.h; setTriMesh + setContactOn,
 * main.C:636-848).  The rigid surfaces: node_count nodes [n][3] with their
 * velocities [n][3] (NULL: at rest), tri_count triangles [t][3] (normal
 * (x1-x0) x (x2-x0), pointing out of the die) and a surface id per triangle
 * (NULL: all 0).  The body must be tetrahedral (the reference's
 * external-face search works on tetrahedron faces, Domain_d.h:243-249); its
 * external faces are found here (SearchExtNodes, Domain_d.C:110-191).  Each
 * step then refreshes the external nodal areas every 10 steps, computes the
 * contact forces after the element forces (first triangle a node is behind
 * and inside of; penalty min(E A/l, 0.2 m/dt^2) with factor
 * wfx_set_contact_params' penalty_factor, damping ratio 0.2, stick/slip
 * Coulomb friction), adds them in calcAccel and moves the surfaces after the
 * position update.  Nodes that leave contact keep their last contact force,
 * as in the reference (its reset is commented out, Contact.C:54). */
int wfx_set_rigid_mesh(wfx_ctx *ctx, int node_count, const double *nodes, const double *node_vel,
                       int tri_count, const int *tris, const int *tri_mesh_id);
/* the contact steps for a caller-driven loop (wfx_run and the fused path
 * run them themselves): Domain_d::CalcExtFaceAreas (Domain_d.C:207-296),
 * CalcContactForces (Contact.C:31-335, after calcElemHourglassForces),
 * TriMesh_d::Move + CalcNormals + UpdatePlaneCoeff (Mesh.h:236-330, after
 * UpdateCorrectionPos) */
int wfx_calc_ext_face_areas(wfx_ctx *ctx, wfx_stream stream);
int wfx_calc_contact_forces(wfx_ctx *ctx, wfx_stream stream);
int wfx_rigid_update(wfx_ctx *ctx, wfx_stream stream);
/* m_contPF ("penaltyFactor", default 0.1), fricCoeffStatic, fricCoeffDynamic
 * (main.C:716-727) */
int wfx_set_contact_params(wfx_ctx *ctx, double penalty_factor, double mu_static, double mu_dynamic);
/* Contact heat conduction: heatCondCoeff and dieTemp of the contact block
 * (main.C:714-719, TriMesh_d::heat_cond and T_const, defaults 0 and 20).
 * With thermal coupling every node in contact takes
 * q = heat_cond * node_area * (die_temp - T) (Contact.C:309), kept until its
 * next contact, and the nodal temperature update adds q / (m cp_T)
 * (Thermal.C:124). */
int wfx_set_contact_heat(wfx_ctx *ctx, double heat_cond, double die_temp);
/* TriMesh_d::node_v: new rigid-node velocities [n][3] (the reference's start-up
 * ramp of the die velocity, Solver_explicit.C:981-992, is a driver's choice) */
int wfx_set_rigid_velocity(wfx_ctx *ctx, const double *node_vel);
/* trimesh->react_force[mesh_id].z of Contact.C:318-321: the sum, in node
 * order, of |normal + damping force| over the nodes in contact with that
 * surface at the last step (synchronises) */
int wfx_contact_force(wfx_ctx *ctx, int mesh_id, double *out);
/* Domain_d::m_max_edot ("maxStrRate", main.C:228; default 1e6): cap of the
 * effective strain rate passed to the tangent modulus (Mechanical.C:1737) */
int wfx_set_max_strain_rate(wfx_ctx *ctx, double edot);
/* per-element initial yield stress (one Material_ per element, mat[e]->sy0) */
int wfx_set_elem_sy0(wfx_ctx *ctx, const double *sy0, int elem_count);
/* Domain_d::AddBCVelNode (Domain_d.C:1057-1061) */
int wfx_add_bc_vel_node(wfx_ctx *ctx, int node, int dim, double val);
/* Domain_d::AllocateBCs (Domain_d.C:1063-1107): uploads the lists */
int wfx_allocate_bcs(wfx_ctx *ctx);
/* Domain_d::SetDT (Domain_d.h) */
int wfx_set_dt(wfx_ctx *ctx, double dt);
/* Domain_d::setCFL + setFixedDt(false) (main.C:263-266): cfl_factor > 0
 * recomputes dt every step as in the BUILD_REMESH CPU solver
 * (Solver_explicit.C:576-600): after the prediction and the Jacobians,
 * dt = cfl_factor * min edge / (sqrt(K/rho[0]) + max |v|); the step's stress
 * update and correction and the next prediction use it.  The dt of
 * wfx_set_dt is then the first prediction's.  0 (default) keeps dt fixed.
 * Runs synchronise at their end to bring dt back to the host. */
int wfx_set_cfl(wfx_ctx *ctx, double cfl_factor);
/* the per-step CFL time step of the current state, as SolveChungHulbert
 * computes it after calcElemJAndDerivatives (Solver_explicit.C:576-600);
 * becomes the dt of the following per-step calls (wfx_run does this itself).
 * Synchronises. */
int wfx_calc_cfl_dt(wfx_ctx *ctx, double *dt, wfx_stream stream);
/* Domain_d::getDT: the current dt (with the per-step CFL time step, the last
 * step's) */
int wfx_get_dt(wfx_ctx *ctx, double *dt);
/* the dt of each step of the last wfx_run / wfx_step_fused / wfx_graph_launch
 * with the per-step CFL time step (count <= its steps) */
int wfx_dt_history(wfx_ctx *ctx, double *out, int count);
/* m_alpha, m_beta, m_gamma (Solver_explicit.C:193-197); defaults rho_b=0.8182 */
int wfx_set_ch(wfx_ctx *ctx, double alpha, double beta, double gamma);
/* Domain_d::m_stab and m_artifvisc (Domain_d.h:139-147, :283-294, main.C:84-120,
 * :293-298) and the 3D hexa hourglass coefficient ("hg_coeff").  Names:
 * alpha_free, alpha_contact, log_factor, J_min, hg_coeff_free,
 * hg_coeff_contact, pspg_scale, p_pspg_bulkfac, av_coeff_div, av_coeff_bulk,
 * artifvisc0, artifvisc1, hg_coeff (all default 0).  With contact on, an
 * element with a node carrying a contact force uses alpha_contact,
 * hg_coeff_contact and the mean of the two viscous pressure terms
 * (calcElemPressure, Mechanical.C:729-799).  artifvisc0/1 add Wilkins'
 * artificial viscosity to the diagonal stresses after the stress update
 * (calcArtificialViscosity, Mechanical.C:1948-1977), on every path.
 * "check_status" (1 by default, no reference counterpart): wfx_run,
 * wfx_step_fused and wfx_graph_launch read the status counters at their end
 * (synchronising) and return WFX_ERR_NONFINITE when they grew; 0 keeps those
 * calls asynchronous and leaves the counters to the query calls. */
int wfx_set_param(wfx_ctx *ctx, const char *name, double value);
/* wfx_get_param also reads K, G, cs0, H (the material's derived moduli) and,
 * read-only, "cluster_count" and "slot_count" (the element clusters of the
 * fused element pass and their (cluster, node) slots; no counterpart). */
int wfx_get_param(wfx_ctx *ctx, const char *name, double *value);
/* nodes whose internal force is summed by wfx_reaction_force (no counterpart
 * in the reference; it has no reaction output): one group */
int wfx_set_reaction_nodes(wfx_ctx *ctx, const int *nodes, int count);
/* several groups (e.g. one per specimen of a batch): group g is
 * nodes[group_offsets[g] .. group_offsets[g+1]); reaction outputs and
 * histories then hold ngroups values per step (no reference counterpart) */
int wfx_set_reaction_groups(wfx_ctx *ctx, const int *nodes, int count, const int *group_offsets,
                            int ngroups);
int wfx_reaction_group_count(wfx_ctx *ctx, int *ngroups);
/* per-element bilinear tangent modulus (one Material_ per element, as the
 * reference's mat[e] pointers allow; used by specimen batches) */
int wfx_set_elem_tangent(wfx_ctx *ctx, const double *Et, int elem_count);
/* Domain_d::calcMinEdgeLength + getMinLength (Domain_d.C:2224-2469): the
 * minimum edge of the tet formed by each element's first four nodes,
 * reduced on the device. */
int wfx_calc_min_edge_length(wfx_ctx *ctx, double *out, wfx_stream stream);

/* ---- init block of SolveChungHulbert (Solver_explicit.C:101-340) --------- */
int wfx_init(wfx_ctx *ctx, wfx_stream stream);

/* ---- per-step API (Solver_explicit.C:524-978 call order) ----------------- */
int wfx_update_prediction(wfx_ctx *ctx, wfx_stream s);          /* Domain_d.C:961-974   */
int wfx_impose_bcv(wfx_ctx *ctx, int dim, wfx_stream s);        /* Domain_d.C:1109-1122 */
int wfx_calc_elem_jacobians(wfx_ctx *ctx, wfx_stream s);        /* Domain_d.C:1701-2044 */
int wfx_calc_elem_vol(wfx_ctx *ctx, wfx_stream s);              /* Mechanical.C:264-293 */
int wfx_calc_elem_density(wfx_ctx *ctx, wfx_stream s);          /* Mechanical.C:295-307 */
int wfx_calc_nodal_vol(wfx_ctx *ctx, wfx_stream s);             /* Mechanical.C:1555-1572 */
int wfx_calc_nodal_mass_from_vol(wfx_ctx *ctx, wfx_stream s);   /* Mechanical.C:1576-1601 */
int wfx_calc_elem_strain_rates(wfx_ctx *ctx, wfx_stream s);     /* Mechanical.C:41-126  */
int wfx_calc_elem_pressure(wfx_ctx *ctx, wfx_stream s);         /* Mechanical.C:691-819 */
int wfx_calc_stress_strain(wfx_ctx *ctx, double dt, wfx_stream s); /* Mechanical.C:1664-1839 */
int wfx_calc_artificial_viscosity(wfx_ctx *ctx, wfx_stream s);  /* Mechanical.C:1948-1977 */
int wfx_calc_elem_forces(wfx_ctx *ctx, wfx_stream s);           /* Mechanical.C:375-481 */
/* Mechanical.C:1842-1943 (3D hexa branch per f90_ver/src/Mechanical.f90:241-303) */
int wfx_calc_elem_hourglass_forces(wfx_ctx *ctx, wfx_stream s);
/* Matrices.C:42-87, plus the non-finite scrub of Solver_explicit.C:777-785 */
int wfx_assembly_forces(wfx_ctx *ctx, wfx_stream s);
int wfx_calc_accel(wfx_ctx *ctx, wfx_stream s);                 /* Mechanical.C:321-341 */
int wfx_impose_bca(wfx_ctx *ctx, int dim, wfx_stream s);        /* Domain_d.C:1124-1134 */
int wfx_update_correction_accvel(wfx_ctx *ctx, wfx_stream s);   /* Domain_d.C:981-997   */
int wfx_update_correction_pos(wfx_ctx *ctx, wfx_stream s);      /* Domain_d.C:1005-1025 */

/* ---- composed ------------------------------------------------------------ */
/* One SolveChungHulbert loop body through the per-step entry points above,
 * in the reference order (Solver_explicit.C:524-978). */
int wfx_step_api(wfx_ctx *ctx, wfx_stream s);
/* nsteps loop bodies; reaction-force history (length nsteps, host memory,
 * may be NULL) is copied back at the end.  Returns WFX_ERR_NONFINITE (after
 * completing the steps, as the reference's loop goes on) when an element
 * Jacobian had detJ <= 0 or an internal force was not finite during the run
 * (see "check_status" under wfx_set_param); the same for wfx_step_fused and
 * wfx_graph_launch. */
int wfx_run(wfx_ctx *ctx, int nsteps, double *reaction_hist, wfx_stream s);
/* nsteps loop bodies of the same sequence regrouped into four fused kernels
 * per step (two element passes, two node passes).  Same state on return as
 * wfx_run, except that the per-step intermediates detJ, dH*, str_rate,
 * rot_rate, strain_pl_incr, f_elem and f_hg are not kept (reading them fails
 * with WFX_ERR_STATE until a per-step API step has run). */
int wfx_step_fused(wfx_ctx *ctx, int nsteps, double *reaction_hist, wfx_stream s);
/* One kernel of the fused step on the current state (no counterpart in the
 * reference; for per-kernel timing and profiling): phase 1 element volumes,
 * 2 nodal sums, 3 element chain, 4 node update (predict: also run the next
 * prediction), 5 sigma = -p I + tau. */
int wfx_fused_phase(wfx_ctx *ctx, int phase, int predict, wfx_stream s);
/* Sum of internal force component dim over each reaction group; out holds
 * one value per group (synchronous).  Histories of wfx_run,
 * wfx_step_fused and wfx_graph_launch are [nsteps][ngroups]. */
int wfx_reaction_force(wfx_ctx *ctx, int dim, double *out, wfx_stream s);

/* ---- launch tuning and graphs (no reference counterpart) ----------------- */
/* Chooses the block size (64/128/256/512) of the fused node passes (phases
 * 2 and 4) on the current mesh: coordinate descent on the time of whole
 * fused steps (max(reps, 10) steps per candidate, best of five rounds, two
 * sweeps), fastest kept (chosen[4], may be NULL; phases 1 and 3 run one
 * 256-thread work-group per element cluster and report 256); the state is
 * saved and restored around the measurement. */
int wfx_autotune(wfx_ctx *ctx, int reps, int *chosen, wfx_stream s);
/* Block sizes of the four fused phases (64/128/256/512 each; the element
 * passes' entries are accepted and not used). */
int wfx_set_block_sizes(wfx_ctx *ctx, const int *block_sizes4);
/* nsteps fused steps as wfx_step_fused runs them, with HIP events recorded on
 * stream s between the four passes: us4[k] = average duration of pass k+1 in
 * microseconds, in the order the passes run.  A measurement call with no
 * reference counterpart; synchronous. */
int wfx_time_fused_phases(wfx_ctx *ctx, int nsteps, double *us4, wfx_stream s);
/* Capture nsteps fused steps (what wfx_step_fused(nsteps) enqueues) into a
 * HIP graph on stream s (a created stream, not NULL); wfx_graph_launch
 * replays it.  A launch fails with WFX_ERR_STATE when parameters, boundary
 * conditions or block sizes changed after the capture. */
int wfx_graph_capture(wfx_ctx *ctx, int nsteps, int with_reaction, wfx_stream s);
int wfx_graph_launch(wfx_ctx *ctx, double *reaction_hist, wfx_stream s);

/* ---- adjoint (no reference counterpart; parameter identification) -------- */
/* Forward: nsteps fused steps from the current state (after wfx_init), as
 * wfx_step_fused, storing the state every checkpoint_every steps; writes the
 * reaction history [nsteps][ngroups] (may be NULL).  Bilinear, Hollomon,
 * Johnson-Cook or GMT material (the rate laws at a uniform or per-element
 * temperature fixed during the run),
 * a fixed time step or the per-step CFL time step, the ANP pressure
 * without stabilisation terms, no contact; thermal coupling only one way (a
 * law that ignores temperature, no thermal expansion: the temperatures
 * follow the mechanics without acting on them).
 * Backward: given dL/dR for that history, recomputes each checkpoint segment
 * and runs the hand-written reverse step over it; grad[3] receives, for the
 * bilinear law, dL/d(sigma_y0) (summed over elements: a uniform yield
 * stress), dL/dEt (summed: a uniform tangent modulus) and dL/d(hg_coeff); for
 * the Hollomon law dL/dK, dL/dm (the plateau stress sigma_y0 held fixed) and
 * dL/d(hg_coeff); for Johnson-Cook dL/dA (summed: a uniform initial yield),
 * dL/dB and dL/d(hg_coeff) (n, C, eps_0, m held fixed; the rate enters
 * through the strain rate of every step); for GMT dL/dC1, dL/dn2 and
 * dL/d(hg_coeff) (the other constants and ranges held fixed).  Optional
 * per-element arrays (length E, may be NULL): each element's share of the
 * first, second and third of these.
 * The state after the backward call is the forward run's final state.
 * Adjoint arithmetic is FP64 in both precisions; in FP32-mixed the forward and
 * recomputation run in FP32-mixed. */
int wfx_adjoint_forward(wfx_ctx *ctx, int nsteps, int checkpoint_every, double *reaction_hist, wfx_stream s);
int wfx_adjoint_backward(wfx_ctx *ctx, const double *reaction_hist_bar, double *grad, double *grad_sy0_elem,
                         double *grad_Et_elem, double *grad_hg_elem, wfx_stream s);
int wfx_adjoint_release(wfx_ctx *ctx);
/* Domain_d::computeEnergies (Mechanical.C:2145-2185), synchronous. */
int wfx_energies(wfx_ctx *ctx, double *ekin, double *deint, wfx_stream s);
/* wfx_energies without any device allocation, for callers that cannot
 * allocate (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The reduction partials live
 * in a static 1 MiB device array, one per GPU, held until the process ends:
 * max(node count, element count) must be at most 16,776,960 (WFX_ERR_ARG
 * beyond).  Calls that overlap on one GPU share that array and race.  Same
 * kernels and results as wfx_energies; synchronous. */
int wfx_energies_noalloc(wfx_ctx *ctx, double *ekin, double *deint, wfx_stream s);
/* number of non-finite internal-force entries scrubbed so far (the scrub of
 * Solver_explicit.C:777-785) */
int wfx_nonfinite_count(wfx_ctx *ctx, long *out);
/* number of element Jacobians found with detJ <= 0 or not finite so far, one
 * count per element and step (the inverted-element check of
 * Solver_explicit.C:639-652, which prints them) */
int wfx_inverted_count(wfx_ctx *ctx, long *out);
/* Bandwidth probe, no reference counterpart: copies bytes (a positive
 * multiple of 16, both pointers 16-byte aligned, device memory) from src to
 * dst on stream s with 16-byte non-temporal accesses; the benchmark times it
 * as the streaming ceiling of the HBM-bound passes. */
int wfx_stream_copy(const void *src, void *dst, long bytes, wfx_stream s);

/* ---- file formats (host side) -------------------------------------------- */
/* LS-DYNA keyword mesh: *NODE and *ELEMENT_SOLID (CreateFromLSDyna,
 * Domain_d.C:1647-1699; its reader is an un-vendored submodule, so this
 * parser is this library's own).  Node ids map to indices in their order under
 * *NODE; solids whose nodes 4..8 repeat node 4 are tetrahedra.  The parse
 * call fills sizes first (x, elnod NULL), then x[3N] and elnod[E*n]. */
int wfx_parse_lsdyna_k(const char *path, int *node_count, int *elem_count, int *nodxelem, double *x, int *elnod);
int wfx_read_lsdyna_k(wfx_ctx *ctx, const char *path);
/* Legacy VTK frame of the current state (VTKWriter.C:365-625 field names:
 * DISP, Velocity, Acceleration, nod_mass, pressure, pl_strain, Vol; plus
 * sigma_y and the cell stress tensor).  Synchronous. */
int wfx_write_vtk(wfx_ctx *ctx, const char *path, wfx_stream s);
/* the same frame, asynchronously: the state is snapshotted on the stream
 * (device-to-device copies ordered with the step), copied to pinned host
 * memory on a library stream and written by a host thread while the caller
 * keeps stepping; two frames in flight at most (a third waits for the
 * oldest).  No reference counterpart (VTKWriter.C writes synchronously). */
int wfx_write_vtk_async(wfx_ctx *ctx, const char *path, wfx_stream stream);
/* waits for every pending asynchronous frame; the first write error */
int wfx_vtk_wait(wfx_ctx *ctx);

/* ---- state transfer (reference layout on the host side, synchronous) ----- */
/* Replace direct access to the Domain_d member arrays (Domain_d.h:841-1043,
 * allocated in SetDimension, Domain_d.C:457-621) and their host mirrors.
 * names: x v a u u_dt prev_a fi fe mdiag voln elnod nodel nodel_loc
 * nodel_offset nodel_count detJ dHdx dHdy dHdz vol vol0 rho rho0 p pl_strain
 * sigma_y sy0 str_rate rot_rate sigma tau eps strain_pl_incr f_elem f_hg;
 * thermal coupling: T q_plheat; contact: contforce ut_prev node_area cfn
 * mesh_in_contact elem_length rigid_x rigid_normal rigid_pplane.
 * Integer fields take int32 buffers, the rest double buffers. */
int wfx_field_count(wfx_ctx *ctx, const char *name, long *count, int *is_int);
int wfx_get_field(wfx_ctx *ctx, const char *name, void *host, long count);
int wfx_set_field(wfx_ctx *ctx, const char *name, const void *host, long count);

#ifdef __cplusplus
}
#endif
#endif
