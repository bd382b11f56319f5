/*
 * wfref — CPU restatement of WeldFormFEM's explicit updated-Lagrangian step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP library in
 * weldformfem_amd/ is checked against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the reference
 * results are compared against or as the timed CPU baseline; the product path
 * never calls it.
 *
 * Every function restates the reference's algorithm in plain C with the
 * reference's own data layout (flat arrays, AoS inside a field) and the same
 * floating-point operation order, so that it can be compiled with
 * -ffp-contract=off and compared bit-for-bit.  Each definition in wfref.c
 * cites the reference file:line it follows.
 *
 * Pinning (what checks this restatement independently of itself):
 *   - the validation/ text logs shipped with the reference (1-element and
 *     4-element runs, with and without hourglass control): golden vectors;
 *   - closed forms (cube Jacobian, patch tests, radial return on a uniaxial
 *     path, ANP under uniform J, Chung-Hulbert free flight);
 *   - invariants (CSR, momentum balance, hourglass dissipation).
 * The reference C++ itself cannot be compiled here (its headers include a
 * header of an un-vendored submodule), so no function is checked against a
 * compiled reference.
 */
#ifndef WFREF_H
#define WFREF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { WFREF_BILINEAR = 0, WFREF_HOLLOMON = 1, WFREF_JOHNSON_COOK = 2, WFREF_GMT = 3 };

typedef struct {
  int model;            /* WFREF_BILINEAR / WFREF_HOLLOMON                  */
  double E, nu, K, G;   /* Elastic_ (Material.cuh:15-34)                    */
  double rho0, cs0;     /* cs0 = sqrt(K/rho0) (main_1_elem_3d.C:113)        */
  double sy0;           /* initial yield                                     */
  double Et, H;         /* bilinear tangent and plastic modulus E*Et/(E-Et)  */
  double Kh, mh, eps0, eps1; /* Hollomon (Material.cuh:90-104)               */
  /* Johnson-Cook (Material.cuh:105-118): A is the per-element sy0[] */
  double B, n, C, eps_0, m, T_m, T_t;
  /* GMT (Material.cuh:271-291) and its clamping ranges */
  double C1, C2, n1, n2, m1, m2, I1, I2;
  double e_min, e_max, er_min, er_max, T_min, T_max;
} wfref_material;

typedef struct {
  int dim, nodxelem, node_count, elem_count, gp_count;
  /* nodal, 3N (AoS inside the field, like Domain_d.h:841-1043) */
  double *x, *v, *a, *u, *u_dt, *prev_a, *fi, *fe;
  double *mdiag, *voln;       /* N */
  /* connectivity */
  int *elnod;                 /* nE */
  int *nodel, *nodel_loc;     /* nE */
  int *nodel_offset, *nodel_count; /* N */
  /* element fields */
  double *detJ, *dHdx, *dHdy, *dHdz;                 /* E, nE */
  double *vol, *vol0, *rho, *rho0, *p, *pl_strain, *sigma_y, *sy0; /* E */
  double *str_rate, *rot_rate, *sigma, *tau, *eps, *strain_pl_incr; /* 6E */
  double *f_elem, *f_hg;                             /* 3nE, [e][a][d] */
  double *temp;       /* element temperature, E (IniTemp 20, main.C:424) */
  /* thermal coupling (Thermal.C, setThermalOn): nodal temperature, plastic
   * heat source, element-node temperature rates */
  int thermal;
  double k_T, cp_T, plheat_frac;
  double exp_T;       /* thermal expansion coefficient (Material_::exp_T, "thermalExp") */
  double *T;          /* N */
  double *q_plheat;   /* E */
  double *dTedt;      /* nE, [e][a] */
  /* rigid-surface penalty contact (Contact.C:31-335, TriMesh_d Mesh.h) */
  int contact;
  double contPF, mu_sta, mu_dyn;   /* m_contPF (Domain_d.h:256), friction */
  int rn, rt;                      /* rigid nodes, rigid triangles */
  double *rnode, *rvel;            /* 3rn */
  int *rtri, *rmesh;               /* 3rt, rt */
  double *rnormal, *rpplane;       /* 3rt, rt */
  int nf;                          /* external faces (count 1), face-list order */
  int *ext_face;                   /* 3nf */
  int *ext_node;                   /* N: 1 on an external face */
  double *node_area, *elem_length; /* N, E */
  double *contforce, *ut_prev;     /* 3N, persistent */
  /* contact heat conduction (heatCondCoeff, dieTemp: main.C:718-719;
   * TriMesh_d::heat_cond, T_const): q_cont_conv of the nodes in contact,
   * kept by a node that leaves contact */
  double heat_cond, T_const;
  double *q_cont_conv;             /* N, persistent */
  int *mesh_in_contact;            /* N */
  double *cfn;                     /* N */
  long step_count;
  double max_edot;    /* strain-rate cap of the tangent (Domain_d.h:824) */
  double cfl;         /* > 0: per-step dt = cfl*minl/(cs+max|v|); 0: fixed */
  /* material (one shared law; per-element initial yield in sy0[]) */
  wfref_material mat;
  /* Dirichlet velocity lists per dimension (AddBCVelNode order) */
  int bc_count[3], bc_cap[3];
  int *bc_nod[3];
  double *bc_val[3];
  /* reaction-force node set */
  int react_count;
  int *react_nod;
  /* time integration */
  double dt, alpha, beta, gamma;
  /* stabilisation (Domain_d.h:283-294 defaults are all zero) */
  double alpha_free, log_factor, J_min, hg_coeff_free, pspg_scale,
         p_pspg_bulkfac, av_coeff_div, av_coeff_bulk;
  double alpha_contact, hg_coeff_contact; /* elements with a node in contact */
  double artifvisc[2];
  double hg_coeff;    /* 3D hexa viscous hourglass coefficient (0 = none) */
  int nonfinite_fi;   /* count of scrubbed non-finite internal forces      */
  int inverted;       /* count of Jacobians with detJ <= 0 or not finite   */
} wfref_domain;

wfref_domain *wfref_create(void);
void wfref_destroy(wfref_domain *d);

/* mesh */
int wfref_add_box_length(wfref_domain *d, const double V[3], const double L[3],
                         double r, int red_int, int tritetra);
int wfref_create_from_mesh(wfref_domain *d, int node_count, int elem_count,
                           int nodxelem, const double *x, const int *elnod);
void wfref_set_nod_elem(wfref_domain *d, const int *elnod);

/* setup */
void wfref_set_density(wfref_domain *d, double rho);
void wfref_set_material(wfref_domain *d, int model, double E, double nu,
                        double rho0, double sy0, double p1, double p2);
void wfref_set_material_jc(wfref_domain *d, double E, double nu, double rho0,
                           double A, double B, double n, double C, double eps_0,
                           double m, double T_m, double T_t);
void wfref_set_material_gmt(wfref_domain *d, double E, double nu, double rho0,
                            double sy0, const double c[8], const double range[6]);
void wfref_set_elem_sy0(wfref_domain *d, const double *sy0);
/* setThermalOn + thermalCond / thermalHeatCap / plHeatFrac (main.C:218,
 * :590-592) */
void wfref_set_thermal(wfref_domain *d, int on, double k_T, double cp_T, double plheat_frac);
void wfref_thermal_calcs(wfref_domain *d);
/* material thermalExp (main.C:577) and calcThermalExpansion (Thermal.C:164-178) */
void wfref_set_thermal_expansion(wfref_domain *d, double exp_T);
void wfref_calc_thermal_expansion(wfref_domain *d);
/* rigid triangle surfaces (TriMesh_d: nodes, node velocities, triangles,
 * mesh id per triangle) for penalty contact on a tetrahedral body; turns
 * contact on, finds the external faces (SearchExtNodes) */
int wfref_set_rigid_mesh(wfref_domain *d, int nn, const double *nodes, const double *vel, int nt,
                         const int *tris, const int *mesh_id);
void wfref_calc_ext_face_areas(wfref_domain *d);
void wfref_calc_elem_length(wfref_domain *d);
void wfref_calc_contact_forces(wfref_domain *d);
void wfref_rigid_update(wfref_domain *d);
double wfref_contact_force(wfref_domain *d, int mesh_id);
/* the hardening law of one element, exposed for the pin tests */
double wfref_eff_strain_rate(const double D[6]);
double wfref_yield_stress(const wfref_domain *d, int e, double ep, double edot);
double wfref_tangent_modulus(const wfref_domain *d, int e, double ep, double edot);
void wfref_add_bc_vel_node(wfref_domain *d, int node, int dim, double val);
void wfref_set_reaction_nodes(wfref_domain *d, int count, const int *nodes);
void wfref_set_ch_default(wfref_domain *d);
void wfref_set_ch(wfref_domain *d, double alpha, double beta, double gamma);
double wfref_calc_min_edge_length(wfref_domain *d);

/* init block of SolveChungHulbert */
void wfref_init(wfref_domain *d);

/* per-step API (Domain_d names) */
void wfref_update_prediction(wfref_domain *d);
void wfref_impose_bcv(wfref_domain *d, int dim);
void wfref_impose_bca(wfref_domain *d, int dim);
void wfref_calc_elem_jacobians(wfref_domain *d);
void wfref_calc_elem_vol(wfref_domain *d);
void wfref_calc_elem_density(wfref_domain *d);
void wfref_calc_nodal_vol(wfref_domain *d);
void wfref_calc_nodal_mass_from_vol(wfref_domain *d);
void wfref_calc_elem_strain_rates(wfref_domain *d);
void wfref_calc_elem_pressure(wfref_domain *d);
void wfref_calc_stress_strain(wfref_domain *d, double dt);
void wfref_calc_artificial_viscosity(wfref_domain *d);
void wfref_calc_elem_forces(wfref_domain *d);
void wfref_calc_elem_hourglass_forces(wfref_domain *d);
void wfref_assembly_forces(wfref_domain *d);
void wfref_scrub_nonfinite_fi(wfref_domain *d);
void wfref_calc_accel(wfref_domain *d);
void wfref_update_correction_accvel(wfref_domain *d);
void wfref_update_correction_pos(wfref_domain *d);

/* one full step in SolveChungHulbert's CPU order; returns reaction force */
double wfref_reaction_force(wfref_domain *d, int dim);
void wfref_step(wfref_domain *d, double *reaction_z);
void wfref_run(wfref_domain *d, int nsteps, double *reaction_z_hist);
void wfref_energies(wfref_domain *d, double *ekin, double *deint);

#ifdef __cplusplus
}
#endif
#endif
