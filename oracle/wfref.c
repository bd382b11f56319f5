/*
 * wfref.c — CPU restatement (oracle) of WeldFormFEM's explicit step.
 * TEST INFRASTRUCTURE ONLY: see wfref.h for who may use it and how it is
 * pinned.  Compile with -ffp-contract=off: the operation order below follows
 * the reference expression by expression (cited file:line, paths relative to
 * the reference repository root).
 */
#include "wfref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RD(p, n) ((p) = (double *)calloc((size_t)(n) > 0 ? (size_t)(n) : 1, sizeof(double)))
#define RI(p, n) ((p) = (int *)calloc((size_t)(n) > 0 ? (size_t)(n) : 1, sizeof(int)))

/* ------------------------------------------------------------------------ */
/* tensor3: full 3x3, same operators as include/common/Tensor3.C            */
/* ------------------------------------------------------------------------ */
typedef struct { double xx, xy, xz, yx, yy, yz, zx, zy, zz; } t3;

/* FromFlatSym (Tensor3.C:75-83): [xx,yy,zz,xy,yz,xz] */
static t3 t3_from_sym(const double *f) {
  t3 r;
  r.xx = f[0]; r.yy = f[1]; r.zz = f[2];
  r.xy = r.yx = f[3];
  r.yz = r.zy = f[4];
  r.xz = r.zx = f[5];
  return r;
}
/* FromFlatAntiSym (Tensor3.C:100-113) */
static t3 t3_from_antisym(const double *f) {
  t3 r;
  r.xx = f[0]; r.yy = f[1]; r.zz = f[2];
  r.xy = f[3]; r.yz = f[4]; r.xz = f[5];
  r.yx = -f[3]; r.zy = -f[4]; r.zx = -f[5];
  return r;
}
/* ToFlatSymPtr (Tensor3.C:120-125): stores the upper triangle */
static void t3_to_sym(const t3 *m, double *f) {
  f[0] = m->xx; f[1] = m->yy; f[2] = m->zz;
  f[3] = m->xy; f[4] = m->yz; f[5] = m->xz;
}
static double t3_trace(const t3 *m) { return m->xx + m->yy + m->zz; }
static t3 t3_identity(void) {
  t3 r;
  r.xx = r.yy = r.zz = 1.;
  r.xy = r.xz = r.yx = r.yz = r.zx = r.zy = 0.;
  return r;
}
/* operator*(tensor3,tensor3) as written in Tensor3.C (operator* after the
 * double3 products): the x column is (A*B) while the y and z columns are
 * (A*B^T).  Kept verbatim because the Jaumann terms depend on it. */
static t3 t3_mul(const t3 *a, const t3 *b) {
  t3 r;
  r.xx = a->xx * b->xx + a->xy * b->yx + a->xz * b->zx;
  r.xy = a->xx * b->yx + a->xy * b->yy + a->xz * b->yz;
  r.xz = a->xx * b->zx + a->xy * b->zy + a->xz * b->zz;
  r.yx = a->yx * b->xx + a->yy * b->yx + a->yz * b->zx;
  r.yy = a->yx * b->yx + a->yy * b->yy + a->yz * b->yz;
  r.yz = a->yx * b->zx + a->yy * b->zy + a->yz * b->zz;
  r.zx = a->zx * b->xx + a->zy * b->yx + a->zz * b->zx;
  r.zy = a->zx * b->yx + a->zy * b->yy + a->zz * b->yz;
  r.zz = a->zx * b->zx + a->zy * b->zy + a->zz * b->zz;
  return r;
}
/* Trans (Tensor3.C) */
static t3 t3_trans(const t3 *m) {
  t3 r;
  r.xx = m->xx; r.yy = m->yy; r.zz = m->zz;
  r.xy = m->yx; r.xz = m->zx;
  r.yx = m->xy; r.yz = m->zy;
  r.zx = m->xz; r.zy = m->yz;
  return r;
}
/* operator*(double,tensor3) and operator*(tensor3,double): m_ij * f */
static t3 t3_scale(const t3 *m, double f) {
  t3 r;
  r.xx = m->xx * f; r.xy = m->xy * f; r.xz = m->xz * f;
  r.yx = m->yx * f; r.yy = m->yy * f; r.yz = m->yz * f;
  r.zx = m->zx * f; r.zy = m->zy * f; r.zz = m->zz * f;
  return r;
}
static t3 t3_add(const t3 *a, const t3 *b) {
  t3 r;
  r.xx = a->xx + b->xx; r.xy = a->xy + b->xy; r.xz = a->xz + b->xz;
  r.yx = a->yx + b->yx; r.yy = a->yy + b->yy; r.yz = a->yz + b->yz;
  r.zx = a->zx + b->zx; r.zy = a->zy + b->zy; r.zz = a->zz + b->zz;
  return r;
}
static t3 t3_sub(const t3 *a, const t3 *b) {
  t3 r;
  r.xx = a->xx - b->xx; r.xy = a->xy - b->xy; r.xz = a->xz - b->xz;
  r.yx = a->yx - b->yx; r.yy = a->yy - b->yy; r.yz = a->yz - b->yz;
  r.zx = a->zx - b->zx; r.zy = a->zy - b->zy; r.zz = a->zz - b->zz;
  return r;
}

/* ------------------------------------------------------------------------ */
/* lifetime                                                                 */
/* ------------------------------------------------------------------------ */
wfref_domain *wfref_create(void) {
  wfref_domain *d = (wfref_domain *)calloc(1, sizeof(wfref_domain));
  wfref_set_ch_default(d);
  d->max_edot = 1.0e6; /* Domain_d.h:824 */
  d->contPF = 0.1;     /* Domain_d.h:256 */
  d->T_const = 20.0;   /* die temperature before "dieTemp" is read (main.C:714) */
  /* m_artifvisc is left uninitialised by the C++ API drivers; zero here */
  return d;
}

static void free_fields(wfref_domain *d) {
  double **dp[] = {&d->x, &d->v, &d->a, &d->u, &d->u_dt, &d->prev_a, &d->fi,
                   &d->fe, &d->mdiag, &d->voln, &d->detJ, &d->dHdx, &d->dHdy,
                   &d->dHdz, &d->vol, &d->vol0, &d->rho, &d->rho0, &d->p,
                   &d->pl_strain, &d->sigma_y, &d->sy0, &d->str_rate,
                   &d->rot_rate, &d->sigma, &d->tau, &d->eps,
                   &d->strain_pl_incr, &d->f_elem, &d->f_hg, &d->temp, &d->T,
                   &d->q_plheat, &d->dTedt, &d->q_cont_conv, &d->node_area, &d->elem_length,
                   &d->contforce, &d->ut_prev, &d->cfn};
  int **ip[] = {&d->elnod, &d->nodel, &d->nodel_loc, &d->nodel_offset,
                &d->nodel_count, &d->ext_face, &d->ext_node, &d->mesh_in_contact};
  for (size_t i = 0; i < sizeof(dp) / sizeof(dp[0]); ++i) { free(*dp[i]); *dp[i] = NULL; }
  for (size_t i = 0; i < sizeof(ip) / sizeof(ip[0]); ++i) { free(*ip[i]); *ip[i] = NULL; }
}

static void free_rigid(wfref_domain *d) {
  free(d->rnode); free(d->rvel); free(d->rtri); free(d->rmesh);
  free(d->rnormal); free(d->rpplane);
  d->rnode = d->rvel = d->rnormal = d->rpplane = NULL;
  d->rtri = d->rmesh = NULL;
  d->rn = d->rt = 0;
}

void wfref_destroy(wfref_domain *d) {
  if (!d) return;
  free_fields(d);
  free_rigid(d);
  for (int k = 0; k < 3; ++k) { free(d->bc_nod[k]); free(d->bc_val[k]); }
  free(d->react_nod);
  free(d);
}

/* SetDimension (Domain_d.C:457-621), every array zero-initialised (the
 * reference relies on fresh pages being zero for several of them). */
static void set_dimension(wfref_domain *d, int node_count, int elem_count) {
  free_fields(d);
  d->node_count = node_count;
  d->elem_count = elem_count;
  int N = node_count, E = elem_count, n = d->nodxelem;
  RD(d->x, 3 * N); RD(d->v, 3 * N); RD(d->a, 3 * N); RD(d->u, 3 * N);
  RD(d->u_dt, 3 * N); RD(d->prev_a, 3 * N); RD(d->fi, 3 * N); RD(d->fe, 3 * N);
  RD(d->mdiag, N); RD(d->voln, N);
  RI(d->elnod, n * E);
  RD(d->detJ, E); RD(d->dHdx, n * E); RD(d->dHdy, n * E); RD(d->dHdz, n * E);
  RD(d->vol, E); RD(d->vol0, E); RD(d->rho, E); RD(d->rho0, E); RD(d->p, E);
  RD(d->pl_strain, E); RD(d->sigma_y, E); RD(d->sy0, E);
  RD(d->str_rate, 6 * E); RD(d->rot_rate, 6 * E); RD(d->sigma, 6 * E);
  RD(d->tau, 6 * E); RD(d->eps, 6 * E); RD(d->strain_pl_incr, 6 * E);
  RD(d->f_elem, 3 * n * E); RD(d->f_hg, 3 * n * E); RD(d->temp, E);
  RD(d->T, N); RD(d->q_plheat, E); RD(d->dTedt, n * E); RD(d->q_cont_conv, N);
  RD(d->node_area, N); RD(d->elem_length, E); RD(d->contforce, 3 * N); RD(d->ut_prev, 3 * N);
  RD(d->cfn, N); RI(d->ext_node, N); RI(d->mesh_in_contact, N);
  free(d->ext_face); d->ext_face = NULL; d->nf = 0; d->contact = 0;
  for (int e = 0; e < E; ++e) d->sy0[e] = d->mat.sy0;
  /* uniform initial temperature, IniTemp default (main.C:424-440) */
  for (int e = 0; e < E; ++e) d->temp[e] = 20.0;
  for (int i = 0; i < N; ++i) d->T[i] = 20.0;
}

/* setNodElem (Domain_d.C:1508-1611): node -> element CSR.  Counts, an
 * exclusive scan for offsets, then a fill in element order so that every
 * node lists its elements in ascending order with the local index. */
void wfref_set_nod_elem(wfref_domain *d, const int *elnod) {
  int N = d->node_count, E = d->elem_count, n = d->nodxelem;
  int *count, *offset;
  RI(count, N);
  RI(offset, N);
  for (int e = 0; e < E; ++e)
    for (int a = 0; a < n; ++a) {
      int g = elnod[e * n + a];
      if (g < N) count[g]++;
    }
  memcpy(d->elnod, elnod, sizeof(int) * (size_t)n * E);
  int tot = 0;
  for (int i = 0; i < N; ++i) { offset[i] = tot; tot += count[i]; }
  free(d->nodel); free(d->nodel_loc);
  RI(d->nodel, tot);
  RI(d->nodel_loc, tot);
  for (int i = 0; i < N; ++i) count[i] = 0;
  for (int e = 0; e < E; ++e)
    for (int a = 0; a < n; ++a) {
      int g = elnod[e * n + a];
      d->nodel[offset[g] + count[g]] = e;
      d->nodel_loc[offset[g] + count[g]] = a;
      count[g]++;
    }
  free(d->nodel_offset); free(d->nodel_count);
  d->nodel_offset = offset;
  d->nodel_count = count;
}

/* AddBoxLength (Domain_d.C:1136-1504), 3D branch.  Coordinates are built by
 * repeated addition of 2r (:1215-1234) and element counts by truncation
 * (:1151-1166).  Hexa local order nb1, nb1+1, nb2+1, nb2, then +nnodz
 * (:1324-1332); six tets per cube (:1383-1388). */
int wfref_add_box_length(wfref_domain *d, const double V[3], const double L[3],
                         double r, int red_int, int tritetra) {
  if (!(L[2] > 0.0) || !red_int) return -1; /* 2D and full integration: out of scope */
  int nel[3];
  nel[0] = (int)(L[0] / (2.0 * r));
  nel[1] = (int)(L[1] / (2.0 * r));
  nel[2] = (int)(L[2] / (2.0 * r));
  d->dim = 3;
  d->gp_count = 1;
  d->nodxelem = tritetra ? 4 : 8;
  int nc = (nel[0] + 1) * (nel[1] + 1) * (nel[2] + 1);
  int ne = nel[0] * nel[1] * nel[2];
  if (tritetra) ne *= 6;
  set_dimension(d, nc, ne);
  double xp, yp, zp = V[2];
  int p = 0;
  for (int k = 0; k < nel[2] + 1; ++k) {
    yp = V[1];
    for (int j = 0; j < nel[1] + 1; ++j) {
      xp = V[0];
      for (int i = 0; i < nel[0] + 1; ++i) {
        d->x[3 * p] = xp; d->x[3 * p + 1] = yp; d->x[3 * p + 2] = zp;
        p++;
        xp = xp + 2.0 * r;
      }
      yp = yp + 2.0 * r;
    }
    zp = zp + 2 * r;
  }
  int *el;
  RI(el, d->nodxelem * ne);
  int ei = 0, nnodz = (nel[0] + 1) * (nel[1] + 1);
  for (int ez = 0; ez < nel[2]; ++ez)
    for (int ey = 0; ey < nel[1]; ++ey)
      for (int ex = 0; ex < nel[0]; ++ex) {
        int nb1 = nnodz * ez + (nel[0] + 1) * ey + ex;
        int nb2 = nnodz * ez + (nel[0] + 1) * (ey + 1) + ex;
        int h[8] = {nb1, nb1 + 1, nb2 + 1, nb2,
                    nb1 + nnodz, nb1 + nnodz + 1, nb2 + nnodz + 1, nb2 + nnodz};
        if (!tritetra) {
          for (int a = 0; a < 8; ++a) el[ei + a] = h[a];
          ei += 8;
        } else {
          static const int T[6][4] = {{0, 1, 3, 5}, {1, 2, 3, 5}, {0, 5, 3, 4},
                                      {4, 5, 3, 7}, {5, 6, 3, 7}, {5, 2, 3, 6}};
          for (int t = 0; t < 6; ++t) {
            for (int a = 0; a < 4; ++a) el[ei + a] = h[T[t][a]];
            ei += 4;
          }
        }
      }
  wfref_set_nod_elem(d, el);
  free(el);
  return 0;
}

/* CreateFromLSDyna (Domain_d.C:1647-1699) without the file reader: node
 * coordinates and element connectivity come in as arrays. */
int wfref_create_from_mesh(wfref_domain *d, int node_count, int elem_count,
                           int nodxelem, const double *x, const int *elnod) {
  if (nodxelem != 4 && nodxelem != 8) return -1;
  d->dim = 3;
  d->gp_count = 1;
  d->nodxelem = nodxelem;
  set_dimension(d, node_count, elem_count);
  memcpy(d->x, x, sizeof(double) * 3 * (size_t)node_count);
  wfref_set_nod_elem(d, elnod);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* setup                                                                    */
/* ------------------------------------------------------------------------ */
/* setDensity (Domain_d.C:952-958) */
void wfref_set_density(wfref_domain *d, double rho) {
  for (int e = 0; e < d->elem_count; ++e) d->rho0[e] = rho;
}

/* Elastic_ (Material.cuh:22-27), Material_::InitHollomon (:90-104).
 * model BILINEAR: p1 = tangent modulus Et; the plastic modulus is
 * H = E*Et/(E-Et) (Material_::Ep, main.C:505) and sigma_y = sy0 + H*ep.
 * model HOLLOMON: p1 = K, p2 = m. */
void wfref_set_material(wfref_domain *d, int model, double E, double nu,
                        double rho0, double sy0, double p1, double p2) {
  wfref_material *m = &d->mat;
  memset(m, 0, sizeof(*m));
  m->model = model;
  m->E = E; m->nu = nu;
  m->K = E / (3.0 * (1.0 - 2.0 * nu));
  m->G = E / (2.0 * (1.0 + nu));
  m->rho0 = rho0;
  m->cs0 = sqrt(m->K / rho0);
  m->sy0 = sy0;
  if (model == WFREF_BILINEAR) {
    m->Et = p1;
    m->H = E * p1 / (E - p1);
  } else {
    m->Kh = p1; m->mh = p2;
    m->eps0 = sy0 / E;
    m->eps1 = pow(sy0 / p1, 1. / p2);
  }
  for (int e = 0; e < d->elem_count; ++e) d->sy0[e] = sy0;
}

/* JohnsonCook(el, A, B, n, C, eps_0, m, T_m, T_t) (Material.cuh:191-203,
 * main.C:535-541: A is yieldStress0, the other seven come from "const").
 * A is kept per element in sy0[] like the bilinear initial yield. */
void wfref_set_material_jc(wfref_domain *d, double E, double nu, double rho0,
                           double A, double B, double n, double C, double eps_0,
                           double m, double T_m, double T_t) {
  wfref_set_material(d, WFREF_BILINEAR, E, nu, rho0, A, 0.0, 0.0);
  wfref_material *mt = &d->mat;
  mt->model = WFREF_JOHNSON_COOK;
  mt->Et = mt->H = 0.0;
  mt->B = B; mt->n = n; mt->C = C; mt->eps_0 = eps_0;
  mt->m = m; mt->T_m = T_m; mt->T_t = T_t;
}

/* GMT(el, n1, n2, C1, C2, m1, m2, I1, I2, e_min, e_max, er_min, er_max,
 * T_min, T_max) (Material.cuh:271-291, main.C:548-553); c[] in that input
 * order, range[] = strain, strain-rate and temperature ranges. */
void wfref_set_material_gmt(wfref_domain *d, double E, double nu, double rho0,
                            double sy0, const double c[8], const double range[6]) {
  wfref_set_material(d, WFREF_BILINEAR, E, nu, rho0, sy0, 0.0, 0.0);
  wfref_material *mt = &d->mat;
  mt->model = WFREF_GMT;
  mt->Et = mt->H = 0.0;
  mt->n1 = c[0]; mt->n2 = c[1]; mt->C1 = c[2]; mt->C2 = c[3];
  mt->m1 = c[4]; mt->m2 = c[5]; mt->I1 = c[6]; mt->I2 = c[7];
  mt->e_min = range[0]; mt->e_max = range[1];
  mt->er_min = range[2]; mt->er_max = range[3];
  mt->T_min = range[4]; mt->T_max = range[5];
}

void wfref_set_elem_sy0(wfref_domain *d, const double *sy0) {
  memcpy(d->sy0, sy0, sizeof(double) * (size_t)d->elem_count);
}

/* AddBCVelNode (Domain_d.C:1057-1061) */
void wfref_add_bc_vel_node(wfref_domain *d, int node, int dim, double val) {
  if (dim < 0 || dim > 2) return;
  if (d->bc_count[dim] == d->bc_cap[dim]) {
    int cap = d->bc_cap[dim] ? 2 * d->bc_cap[dim] : 64;
    d->bc_nod[dim] = (int *)realloc(d->bc_nod[dim], sizeof(int) * (size_t)cap);
    d->bc_val[dim] = (double *)realloc(d->bc_val[dim], sizeof(double) * (size_t)cap);
    d->bc_cap[dim] = cap;
  }
  d->bc_nod[dim][d->bc_count[dim]] = node;
  d->bc_val[dim][d->bc_count[dim]] = val;
  d->bc_count[dim]++;
}

void wfref_set_reaction_nodes(wfref_domain *d, int count, const int *nodes) {
  free(d->react_nod);
  d->react_count = count;
  d->react_nod = (int *)malloc(sizeof(int) * (size_t)(count > 0 ? count : 1));
  memcpy(d->react_nod, nodes, sizeof(int) * (size_t)count);
}

/* Chung-Hulbert constants (Solver_explicit.C:193-197), rho_b = 0.8182 */
void wfref_set_ch_default(wfref_domain *d) {
  double rho_b = 0.818200;
  d->alpha = (2.0 * rho_b - 1.0) / (1.0 + rho_b);
  d->beta = (5.0 - 3.0 * rho_b) / ((1.0 + rho_b) * (1.0 + rho_b) * (2.0 - rho_b));
  d->gamma = 1.5 - d->alpha;
}
void wfref_set_ch(wfref_domain *d, double alpha, double beta, double gamma) {
  d->alpha = alpha; d->beta = beta; d->gamma = gamma;
}

/* calcMinEdgeLength (Domain_d.C:2224-2260): every 3D element is taken as the
 * tet of its first four nodes; minimum of its six edge lengths. */
double wfref_calc_min_edge_length(wfref_domain *d) {
  double min_len = 1.0e6;
  int n = d->nodxelem;
  for (int e = 0; e < d->elem_count; ++e) {
    const int *en = d->elnod + n * e;
    double P[4][3];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) P[k][c] = d->x[3 * en[k] + c];
    static const int EP[6][2] = {{1, 0}, {2, 0}, {3, 0}, {2, 1}, {3, 1}, {3, 2}};
    for (int i = 0; i < 6; ++i) {
      double ex = P[EP[i][0]][0] - P[EP[i][1]][0];
      double ey = P[EP[i][0]][1] - P[EP[i][1]][1];
      double ez = P[EP[i][0]][2] - P[EP[i][1]][2];
      double len = sqrt(ex * ex + ey * ey + ez * ez);
      if (len < min_len) min_len = len;
    }
  }
  return min_len;
}

/* ------------------------------------------------------------------------ */
/* per-step functions                                                       */
/* ------------------------------------------------------------------------ */

/* UpdatePrediction (Domain_d.C:961-974) */
void wfref_update_prediction(wfref_domain *d) {
  int N3 = 3 * d->node_count;
  double dt = d->dt;
#pragma omp parallel for
  for (int ig = 0; ig < N3; ++ig) {
    d->u_dt[ig] = dt * (d->v[ig] + (0.5 - d->beta) * dt * d->prev_a[ig]);
    d->v[ig] += (1.0 - d->gamma) * dt * d->prev_a[ig];
  }
}

/* ImposeBCV (Domain_d.C:1109-1122): list order, the last entry wins */
void wfref_impose_bcv(wfref_domain *d, int dim) {
  for (int i = 0; i < d->bc_count[dim]; ++i)
    d->v[3 * d->bc_nod[dim][i] + dim] = d->bc_val[dim][i];
}
/* ImposeBCA (Domain_d.C:1124-1134) */
void wfref_impose_bca(wfref_domain *d, int dim) {
  for (int i = 0; i < d->bc_count[dim]; ++i)
    d->a[3 * d->bc_nod[dim][i] + dim] = 0.0;
}

/* adjugate = transposed cofactor, not divided by det (Matrix.h:694-727) */
static void adj33(const double J[3][3], double A[3][3]) {
  double C[3][3];
  C[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
  C[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
  C[0][2] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  C[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
  C[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
  C[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
  C[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
  C[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
  C[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = C[j][i];
}
/* calcDet, 3x3 (Matrix.h:593-618) */
static double det33(const double J[3][3]) {
  return J[0][0] * J[1][1] * J[2][2] - J[0][0] * J[1][2] * J[2][1] -
         J[0][1] * J[1][0] * J[2][2] + J[0][1] * J[1][2] * J[2][0] +
         J[0][2] * J[1][0] * J[2][1] - J[0][2] * J[1][1] * J[2][0];
}

/* calcElemJAndDerivatives (Domain_d.C:1701-2044): hexa rows :1822-1831,
 * derivatives :1843-1850; tet :1855-1887; detJ :1894; store :2011-2029.
 * Stores dH_detJ = adj(J) * dN/dxi, i.e. dN/dx * detJ. */
void wfref_calc_elem_jacobians(wfref_domain *d) {
  int n = d->nodxelem, E = d->elem_count;
#pragma omp parallel for
  for (int e = 0; e < E; ++e) {
    double X[8][3];
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < 3; ++c) X[a][c] = d->x[3 * d->elnod[e * n + a] + c];
    double J[3][3], A[3][3], dH[3][8];
    if (n == 8) {
      for (int c = 0; c < 3; ++c) {
        J[0][c] = 0.125 * (-X[0][c] + X[1][c] + X[2][c] - X[3][c] - X[4][c] + X[5][c] + X[6][c] - X[7][c]);
        J[1][c] = 0.125 * (-X[0][c] - X[1][c] + X[2][c] + X[3][c] - X[4][c] - X[5][c] + X[6][c] + X[7][c]);
        J[2][c] = 0.125 * (-X[0][c] - X[1][c] - X[2][c] - X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
      }
      adj33(J, A);
      for (int c = 0; c < 3; ++c) {
        dH[c][0] = 0.125 * (-A[c][0] - A[c][1] - A[c][2]);
        dH[c][1] = 0.125 * (A[c][0] - A[c][1] - A[c][2]);
        dH[c][2] = 0.125 * (A[c][0] + A[c][1] - A[c][2]);
        dH[c][3] = 0.125 * (-A[c][0] + A[c][1] - A[c][2]);
        dH[c][4] = 0.125 * (-A[c][0] - A[c][1] + A[c][2]);
        dH[c][5] = 0.125 * (A[c][0] - A[c][1] + A[c][2]);
        dH[c][6] = 0.125 * (A[c][0] + A[c][1] + A[c][2]);
        dH[c][7] = 0.125 * (-A[c][0] + A[c][1] + A[c][2]);
      }
    } else {
      for (int c = 0; c < 3; ++c) {
        J[0][c] = X[1][c] - X[0][c];
        J[1][c] = X[2][c] - X[0][c];
        J[2][c] = X[3][c] - X[0][c];
      }
      adj33(J, A);
      for (int c = 0; c < 3; ++c) {
        dH[c][0] = -A[c][0] - A[c][1] - A[c][2];
        dH[c][1] = A[c][0];
        dH[c][2] = A[c][1];
        dH[c][3] = A[c][2];
      }
    }
    d->detJ[e] = det33(J);
    /* inverted or degenerate (the check of Solver_explicit.C:639-652) */
    if (!(d->detJ[e] > 0.0)) {
#pragma omp atomic
      d->inverted++;
    }
    for (int a = 0; a < n; ++a) {
      d->dHdx[e * n + a] = dH[0][a];
      d->dHdy[e * n + a] = dH[1][a];
      d->dHdz[e * n + a] = dH[2][a];
    }
  }
}

static double vol_weight(int n) { return n == 4 ? 1.0 / 6.0 : 8.0; }

/* CalcElemVol (Mechanical.C:264-293): vol = detJ * w * 1.0 */
void wfref_calc_elem_vol(wfref_domain *d) {
  double w = vol_weight(d->nodxelem);
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    d->vol[e] = 0.0;
    d->vol[e] += d->detJ[e] * w * 1.0;
  }
}

/* calcElemDensity (Mechanical.C:295-307) */
void wfref_calc_elem_density(wfref_domain *d) {
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e)
    d->rho[e] = d->rho0[e] * d->vol0[e] / d->vol[e];
}

/* CalcNodalVol (Mechanical.C:1555-1572; serial in the reference, every node
 * is independent so the loop is parallel here with identical results) */
void wfref_calc_nodal_vol(wfref_domain *d) {
#pragma omp parallel for
  for (int i = 0; i < d->node_count; ++i) {
    double s = 0.0;
    for (int k = 0; k < d->nodel_count[i]; ++k) s += d->vol[d->nodel[d->nodel_offset[i] + k]];
    s /= d->nodxelem;
    d->voln[i] = s;
  }
}

/* CalcNodalMassFromVol (Mechanical.C:1576-1601) */
void wfref_calc_nodal_mass_from_vol(wfref_domain *d) {
#pragma omp parallel for
  for (int i = 0; i < d->node_count; ++i) {
    double mass = 0.0, f = 1.0;
    for (int k = 0; k < d->nodel_count[i]; ++k) {
      int eg = d->nodel[d->nodel_offset[i] + k];
      mass += f * d->rho[eg] * d->voln[i] / d->nodel_count[i];
    }
    d->mdiag[i] = mass;
  }
}

/* calcElemStrainRates (Mechanical.C:41-126) */
void wfref_calc_elem_strain_rates(wfref_domain *d) {
  int n = d->nodxelem;
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    double S[3][3] = {{0}}, W[3][3] = {{0}};
    double f = 1.0 / d->detJ[e];
    for (int a = 0; a < n; ++a) {
      int g = d->elnod[e * n + a];
      double vx = d->v[3 * g], vy = d->v[3 * g + 1], vz = d->v[3 * g + 2];
      double h0 = d->dHdx[e * n + a], h1 = d->dHdy[e * n + a], h2 = d->dHdz[e * n + a];
      S[0][0] = S[0][0] + h0 * f * vx;
      S[1][1] = S[1][1] + h1 * f * vy;
      S[2][2] = S[2][2] + h2 * f * vz;
      S[0][1] = S[0][1] + f * (h1 * vx + h0 * vy);
      W[0][1] = W[0][1] + f * (h1 * vx - h0 * vy);
      S[1][2] = S[1][2] + f * (h2 * vy + h1 * vz);
      S[0][2] = S[0][2] + f * (h2 * vx + h0 * vz);
      W[1][2] = W[1][2] + f * (h2 * vy - h1 * vz);
      W[0][2] = W[0][2] + f * (h2 * vx - h0 * vz);
    }
    S[0][1] = S[0][1] * 0.5; S[0][2] = S[0][2] * 0.5; S[1][2] = S[1][2] * 0.5;
    W[0][1] = W[0][1] * 0.5; W[0][2] = W[0][2] * 0.5; W[1][2] = W[1][2] * 0.5;
    double *sr = d->str_rate + 6 * e, *rr = d->rot_rate + 6 * e;
    sr[0] = S[0][0]; sr[1] = S[1][1]; sr[2] = S[2][2];
    sr[3] = S[0][1]; sr[4] = S[1][2]; sr[5] = S[0][2];
    rr[0] = 0.0; rr[1] = 0.0; rr[2] = 0.0;
    rr[3] = W[0][1]; rr[4] = W[1][2]; rr[5] = W[0][2];
  }
}

static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* calcElemPressure (Mechanical.C:691-819), press_algorithm 0.
 * Node part :698-705, element part :720-815.  With the default (all zero)
 * stabilisation this is p = -K * (mean over element nodes of V_n/V_n0 - 1).
 * With contact on, an element with a node whose contact force (the previous
 * step's, contforce is never reset) has a positive squared length uses
 * alpha_contact and hg_coeff_contact and p_q = (q1 + q2) / 2 (:729-745,
 * :757, :779, :792-799). */
void wfref_calc_elem_pressure(wfref_domain *d) {
  int N = d->node_count, n = d->nodxelem;
  double *voln0 = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  double *voln = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    voln0[i] = voln[i] = 0.0;
    for (int k = 0; k < d->nodel_count[i]; ++k) {
      int e = d->nodel[d->nodel_offset[i] + k];
      voln0[i] += d->vol0[e];
      voln[i] += d->vol[e];
    }
  }
  double K = d->mat.K;
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    double rho_e = d->rho[e];
    double vol0 = d->vol0[e], vol1 = d->vol[e];
    double J_local = vol1 / vol0;
    double h = pow(vol1, 1.0 / 3.0);
    int is_contact = 0;
    if (d->contact) {
      for (int a = 0; a < n; ++a) {
        const double *cf = d->contforce + 3 * d->elnod[e * n + a];
        if (cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2] > 0) {
          is_contact = 1;
          break;
        }
      }
    }
    double J_avg = 0.0;
    for (int a = 0; a < n; ++a) {
      int nid = d->elnod[e * n + a];
      J_avg += voln[nid] / voln0[nid];
    }
    J_avg /= n;
    double alpha = is_contact ? d->alpha_contact : d->alpha_free;
    double J_bar = alpha * J_local + (1 - alpha) * J_avg;
    if (J_bar < d->J_min) J_bar = 0.2;
    double p_physical = -K * (d->log_factor * log(J_bar) + (1.0 - d->log_factor) * (J_bar - 1.0));
    double c = sqrt(K / rho_e);
    double tau = h / (2.0 * c);
    double div_v = 0.0;
    for (int a = 0; a < n; ++a) {
      int nid = d->elnod[e * n + a];
      div_v += d->dHdx[e * n + a] * d->v[3 * nid] + d->dHdy[e * n + a] * d->v[3 * nid + 1] +
               d->dHdz[e * n + a] * d->v[3 * nid + 2];
    }
    double p_pspg = 0.0;
    double p_hg = (is_contact ? d->hg_coeff_contact : d->hg_coeff_free) * K * fabs(J_local - J_avg);
    double p_q = 0.0;
    if (div_v < 0.0) {
      p_pspg = dmin(d->pspg_scale * tau * div_v * K, d->p_pspg_bulkfac * K);
      double q1 = d->av_coeff_div * rho_e * h * c * (-div_v);
      double delta_J = 1.0 - J_local;
      double q2 = d->av_coeff_bulk * K * delta_J;
      p_q = is_contact ? 0.5 * (q1 + q2) : dmax(q1, q2);
    }
    d->p[e] = p_physical + p_pspg + p_hg + p_q;
  }
  free(voln0);
  free(voln);
}

/* effective strain rate of CalcStressStrain (Mechanical.C:1700-1705) */
double wfref_eff_strain_rate(const double D[6]) {
  double xx = D[0], yy = D[1], zz = D[2], xy = D[3], yz = D[4], zx = D[5];
  return sqrt(0.5 * ((xx - yy) * (xx - yy) + (yy - zz) * (yy - zz) + (zz - xx) * (zz - xx)) +
              3.0 * (xy * xy + yz * yz + zx * zx));
}

/* the clamps of CalcGMTYieldStress / CalcGMTTangentModulus
 * (Material.cuh:427-435, :451-459) */
static double clampr(double v, double lo, double hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}

/* The temperature the hardening law of element e reads.  Without thermal
 * coupling: temp[e].  With it: the mean of the element's nodal
 * temperatures (the reference reads the nodal array T at the element index,
 * Mechanical.C:1725-1729, which is defined only for e < node count). */
static double law_temperature(const wfref_domain *d, int e) {
  if (!d->thermal) return d->temp[e];
  int n = d->nodxelem;
  double t = 0.0;
  for (int a = 0; a < n; ++a) t += d->T[d->elnod[e * n + a]];
  return t / n;
}

/* Calc*YieldStress (Material.cuh:353-364, :378-388, :412-441) of element e
 * at plastic strain ep and effective strain rate edot; the element
 * temperature is law_temperature(d, e) (the reference reads T[e], :1725-1729 of
 * Mechanical.C).  BILINEAR is this library's definition (the reference leaves
 * it undefined). */
double wfref_yield_stress(const wfref_domain *d, int e, double ep, double edot) {
  const wfref_material *m = &d->mat;
  if (m->model == WFREF_HOLLOMON) {
    if (ep + m->eps0 > m->eps1) return m->Kh * pow(ep + m->eps0, m->mh);
    return m->sy0;
  }
  if (m->model == WFREF_JOHNSON_COOK) {
    double T_h = (law_temperature(d, e) - m->T_t) / (m->T_m - m->T_t);
    double sr = edot;
    if (edot == 0.0) sr = 1.e-5;
    return (d->sy0[e] + m->B * pow(ep, m->n)) * (1.0 + m->C * log(sr / m->eps_0)) *
           (1.0 - pow(T_h, m->m));
  }
  if (m->model == WFREF_GMT) {
    double es = clampr(ep, m->e_min, m->e_max);
    double er = clampr(edot, m->er_min, m->er_max);
    double T = clampr(law_temperature(d, e), m->T_min, m->T_max);
    return m->C1 * exp(m->C2 * T) * pow(es, m->n1 * T + m->n2) * exp((m->I1 * T + m->I2) / es) *
           pow(er, m->m1 * T + m->m2);
  }
  return d->sy0[e] + m->H * ep;
}

/* Calc*TangentModulus (Material.cuh:389-395, :397-411, :444-473), with the
 * strain rate already capped by max_edot (Mechanical.C:1737). */
double wfref_tangent_modulus(const wfref_domain *d, int e, double ep, double edot) {
  const wfref_material *m = &d->mat;
  if (m->model == WFREF_HOLLOMON) {
    if (ep + m->eps0 > m->eps1) return m->Kh * m->mh * pow(ep + m->eps0, (m->mh - 1.0));
    return 0.;
  }
  if (m->model == WFREF_JOHNSON_COOK) {
    double T_h = (law_temperature(d, e) - m->T_t) / (m->T_m - m->T_t);
    if (ep > 0.)
      return m->n * m->B * pow(ep, m->n - 1.) * (1.0 + m->C * log(edot / m->eps_0)) *
             (1.0 - pow(T_h, m->m));
    return m->E * 0.1;
  }
  if (m->model == WFREF_GMT) {
    double es = clampr(ep, m->e_min, m->e_max);
    double er = clampr(edot, m->er_min, m->er_max);
    double T = clampr(law_temperature(d, e), m->T_min, m->T_max);
    return m->C1 * exp(m->C2 * T) * pow(er, m->m1 * T + m->m2) * pow(es, T * m->n1 + m->n2 - 2.0) *
           (-m->I1 * T - m->I2 + es * (m->n1 * T + m->n2)) * exp((m->I1 * T + m->I2) / es);
  }
  return m->H;
}

/* CalcStressStrain (Mechanical.C:1664-1839): Jaumann update of the
 * deviatoric stress (:1690-1698), trial state (:1710-1721), J2 radial return
 * (:1723-1762), total stress (:1766), plastic strain increment (:1784-1796),
 * strain update and stores (:1820-1829). */
void wfref_calc_stress_strain(wfref_domain *d, double dt) {
  double G = d->mat.G;
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    t3 tau = t3_from_sym(d->tau + 6 * e);
    t3 D = t3_from_sym(d->str_rate + 6 * e);
    t3 W = t3_from_antisym(d->rot_rate + 6 * e);
    t3 Eps = t3_from_sym(d->eps + 6 * e);
    t3 I = t3_identity();
    t3 Wt = t3_trans(&W);
    t3 SRT = t3_mul(&tau, &Wt);
    t3 RS = t3_mul(&W, &tau);
    t3 cI = t3_scale(&I, 1.0 / 3.0 * t3_trace(&D));
    t3 devD = t3_sub(&D, &cI);
    t3 A = t3_scale(&devD, 2.0 * G);
    t3 B = t3_add(&A, &SRT);
    t3 C = t3_add(&B, &RS);
    t3 Dt = t3_scale(&C, dt);
    tau = t3_add(&tau, &Dt);

    double p = d->p[e];
    t3 pI = t3_scale(&I, -p);
    t3 sig_tr = t3_add(&pI, &tau);
    t3 trI = t3_scale(&I, (1.0 / 3.0) * t3_trace(&sig_tr));
    t3 s = t3_sub(&sig_tr, &trI);
    double J2 = 0.5 * (s.xx * s.xx + 2.0 * s.xy * s.xy + 2.0 * s.xz * s.xz + s.yy * s.yy +
                       2.0 * s.yz * s.yz + s.zz * s.zz);
    double sig_trial = sqrt(3.0 * J2);
    double edot = wfref_eff_strain_rate(d->str_rate + 6 * e);
    d->sigma_y[e] = wfref_yield_stress(d, e, d->pl_strain[e], edot);
    edot = dmin(edot, d->max_edot);
    double dep = 0.0;
    if (d->sigma_y[e] < sig_trial) {
      double H = wfref_tangent_modulus(d, e, d->pl_strain[e], edot);
      double dgamma = (sig_trial - d->sigma_y[e]) / (3.0 * G + H);
      double factor = 1.0 - (3.0 * G * dgamma) / sig_trial;
      tau = t3_scale(&s, factor);
      d->pl_strain[e] += dgamma;
      dep = dgamma;
    }
    t3 sig = t3_add(&pI, &tau);
    double *pi = d->strain_pl_incr + 6 * e;
    if (d->thermal) d->q_plheat[e] = 0.0;
    if (dep > 0.0) {
      double f = dep / d->sigma_y[e];
      pi[0] = f * (sig.xx - 0.5 * (sig.yy + sig.zz));
      pi[1] = f * (sig.yy - 0.5 * (sig.xx + sig.zz));
      pi[2] = f * (sig.zz - 0.5 * (sig.xx + sig.yy));
      pi[3] = 1.5 * f * (sig.xy);
      pi[4] = 1.5 * f * (sig.yz);
      pi[5] = 1.5 * f * (sig.xz);
      if (d->thermal) {
        /* plastic work rate (Mechanical.C:1798-1810): depdt = (1/dt) incr */
        double r = 1. / dt;
        d->q_plheat[e] = d->plheat_frac * (sig.xx * (r * pi[0]) + 2.0 * sig.xy * (r * pi[3]) +
                                           2.0 * sig.xz * (r * pi[5]) + sig.yy * (r * pi[1]) +
                                           2.0 * sig.yz * (r * pi[4]) + sig.zz * (r * pi[2]));
      }
    } else {
      for (int k = 0; k < 6; ++k) pi[k] = 0.0;
    }
    t3 Ddt = t3_scale(&D, dt);
    Eps = t3_add(&Eps, &Ddt);
    t3_to_sym(&sig, d->sigma + 6 * e);
    t3_to_sym(&tau, d->tau + 6 * e);
    t3_to_sym(&Eps, d->eps + 6 * e);
  }
}

/* calcArtificialViscosity (Mechanical.C:1948-1977) */
void wfref_calc_artificial_viscosity(wfref_domain *d) {
  double al = d->artifvisc[0], be = d->artifvisc[1], q_max = 1e9;
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    double c = sqrt(d->mat.K / d->rho[e]);
    const double *sr = d->str_rate + 6 * e;
    double eps_v = sr[0] + sr[1] + sr[2];
    if (fabs(eps_v) > 1e-12) {
      double l = pow(d->vol[e], 1.0 / 3.0);
      double q = al * c * fabs(eps_v) * l + be * pow(eps_v * l, 2);
      q = dmin(q, q_max);
      double qs = (eps_v > 0) ? -q : q;
      d->sigma[6 * e + 0] += qs;
      d->sigma[6 * e + 1] += qs;
      d->sigma[6 * e + 2] += qs;
    }
  }
}

/* calcElemForces (Mechanical.C:375-481), 3D branch :448-457; w = 8 (hexa)
 * or 1/6 (tet); f[e][a][d]. */
void wfref_calc_elem_forces(wfref_domain *d) {
  int n = d->nodxelem;
  double w = vol_weight(n);
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    double *f = d->f_elem + 3 * n * e;
    const double *s = d->sigma + 6 * e; /* xx yy zz xy yz xz */
    for (int k = 0; k < 3 * n; ++k) f[k] = 0.0;
    for (int a = 0; a < n; ++a) {
      double h0 = d->dHdx[e * n + a], h1 = d->dHdy[e * n + a], h2 = d->dHdz[e * n + a];
      f[3 * a + 0] += h0 * s[0] * 1.0;
      f[3 * a + 1] += h1 * s[1] * 1.0;
      f[3 * a + 2] += h2 * s[2] * 1.0;
      f[3 * a + 0] += h1 * s[3] + h2 * s[5];
      f[3 * a + 1] += h0 * s[3] + h2 * s[4];
      f[3 * a + 2] += h1 * s[4] + h0 * s[5];
    }
    for (int k = 0; k < 3 * n; ++k) f[k] *= w;
  }
}

/* 3D hexa viscous hourglass forces.  The reference C++ leaves this branch
 * empty (Mechanical.C:1845, :1939-1942); the formula is the Fortran one
 * (f90_ver/src/Mechanical.f90:241-303): Goudreau/Flanagan-Belytschko base
 * vectors (:259-263, scaled by 8 to +-1), h(d,j) = sum_a v(a,d)*G_j(a),
 * f(a,d) = -c_h * sum_j h(d,j)*G_j(a), c_h = coeff*vol^0.6666666*rho*0.25*cs0
 * (:299).  Assembly subtracts it (f90_ver/src/Matrices.f90:641). */
static const double HG_GAMMA[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1},
                                      {1, -1, -1, 1, -1, 1, 1, -1},
                                      {1, -1, 1, -1, 1, -1, 1, -1},
                                      {-1, 1, -1, 1, 1, -1, 1, -1}};
void wfref_calc_elem_hourglass_forces(wfref_domain *d) {
  int n = d->nodxelem;
  if (n != 8 || d->hg_coeff == 0.0) return; /* f_hg stays zero */
  double cs0 = d->mat.cs0, coeff = d->hg_coeff;
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    double V[8][3];
    for (int a = 0; a < 8; ++a) {
      int g = d->elnod[8 * e + a];
      for (int c = 0; c < 3; ++c) V[a][c] = d->v[3 * g + c];
    }
    double hm[3][4] = {{0}};
    for (int j = 0; j < 4; ++j)
      for (int a = 0; a < 8; ++a)
        for (int c = 0; c < 3; ++c) hm[c][j] = hm[c][j] + V[a][c] * HG_GAMMA[j][a];
    double *f = d->f_hg + 24 * e;
    for (int k = 0; k < 24; ++k) f[k] = 0.0;
    for (int a = 0; a < 8; ++a)
      for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 3; ++c) f[3 * a + c] = f[3 * a + c] - hm[c][j] * HG_GAMMA[j][a];
    double c_h = coeff * pow(d->vol[e], 0.6666666) * d->rho[e] * 0.25 * cs0;
    for (int k = 0; k < 24; ++k) f[k] = f[k] * c_h;
  }
}

/* assemblyForces (Matrices.C:42-87): per node, sum f_elem over the CSR in
 * order, then subtract f_hg in a second pass over the CSR. */
void wfref_assembly_forces(wfref_domain *d) {
  int n = d->nodxelem;
#pragma omp parallel for
  for (int i = 0; i < d->node_count; ++i) {
    double f[3] = {0.0, 0.0, 0.0};
    int o = d->nodel_offset[i], c = d->nodel_count[i];
    for (int k = 0; k < c; ++k) {
      int off = d->nodel[o + k] * n * 3 + d->nodel_loc[o + k] * 3;
      for (int q = 0; q < 3; ++q) f[q] += d->f_elem[off + q];
    }
    for (int k = 0; k < c; ++k) {
      int off = d->nodel[o + k] * n * 3 + d->nodel_loc[o + k] * 3;
      for (int q = 0; q < 3; ++q) f[q] -= d->f_hg[off + q];
    }
    for (int q = 0; q < 3; ++q) d->fi[3 * i + q] = f[q];
  }
}

/* non-finite internal force scrub (Solver_explicit.C:777-785) */
void wfref_scrub_nonfinite_fi(wfref_domain *d) {
  for (int i = 0; i < 3 * d->node_count; ++i)
    if (!isfinite(d->fi[i])) { d->fi[i] = 0.0; d->nonfinite_fi++; }
}

/* calcAccel (Mechanical.C:321-341), no contact */
void wfref_calc_accel(wfref_domain *d) {
#pragma omp parallel for
  for (int i = 0; i < d->node_count; ++i)
    for (int c = 0; c < 3; ++c)
      d->a[3 * i + c] = (d->fe[3 * i + c] - d->fi[3 * i + c]) / d->mdiag[i];
  if (d->contact) /* Mechanical.C:330-335 */
#pragma omp parallel for
    for (int i = 0; i < d->node_count; ++i)
      for (int c = 0; c < 3; ++c) d->a[3 * i + c] += d->contforce[3 * i + c] / d->mdiag[i];
}

/* UpdateCorrectionAccVel (Domain_d.C:981-997) */
void wfref_update_correction_accvel(wfref_domain *d) {
  double f = 1.0 / (1.0 - d->alpha);
  int N3 = 3 * d->node_count;
#pragma omp parallel for
  for (int ig = 0; ig < N3; ++ig) {
    d->a[ig] = f * (d->a[ig] - d->alpha * d->prev_a[ig]);
    d->v[ig] += d->gamma * d->dt * d->a[ig];
  }
}

/* UpdateCorrectionPos (Domain_d.C:1005-1025) */
void wfref_update_correction_pos(wfref_domain *d) {
  int N3 = 3 * d->node_count;
  double dt = d->dt;
#pragma omp parallel for
  for (int ig = 0; ig < N3; ++ig) {
    d->u_dt[ig] += d->beta * dt * dt * d->a[ig];
    d->x[ig] += d->u_dt[ig];
    d->prev_a[ig] = d->a[ig];
    d->u[ig] += d->u_dt[ig];
  }
}

/* ------------------------------------------------------------------------ */
/* init block and step loop                                                 */
/* ------------------------------------------------------------------------ */

/* SolveChungHulbert init, CPU branch (Solver_explicit.C:101-340):
 * InitValues (Domain_d.C:393-416), v = a = u = 0 and ImposeBCV (:176-190),
 * Jacobians, initial and current volume, density, nodal volume and mass
 * (:253-268).  prev_a and u_dt are zero here (never set in the reference). */
void wfref_init(wfref_domain *d) {
  for (int e = 0; e < d->elem_count; ++e) {
    d->pl_strain[e] = 0.0;
    d->sigma_y[e] = d->sy0[e];
  }
  for (int c = 0; c < 3; ++c) {
    for (int i = 0; i < 3 * d->node_count; ++i) d->v[i] = d->a[i] = d->u[i] = 0.0;
    wfref_impose_bcv(d, c);
  }
  wfref_calc_elem_jacobians(d);
  d->inverted = 0; /* counted from the first step on */
  wfref_calc_elem_vol(d);
  memcpy(d->vol0, d->vol, sizeof(double) * (size_t)d->elem_count);
  wfref_calc_elem_vol(d);
  wfref_calc_elem_density(d);
  wfref_calc_nodal_vol(d);
  wfref_calc_nodal_mass_from_vol(d);
  if (d->contact) {
    /* calcMinEdgeLength before the solve (main.C:862) sets m_elem_length */
    wfref_calc_elem_length(d);
    for (int i = 0; i < 3 * d->node_count; ++i) d->contforce[i] = d->ut_prev[i] = 0.0;
    d->step_count = 0;
  }
}

double wfref_reaction_force(wfref_domain *d, int dim) {
  double r = 0.0;
  for (int k = 0; k < d->react_count; ++k) r += d->fi[3 * d->react_nod[k] + dim];
  return r;
}

/* setThermalOn and the thermal constants (the nodal temperatures stay as
 * set: 20 at mesh creation, setTemp through the "T" field) */
void wfref_set_thermal(wfref_domain *d, int on, double k_T, double cp_T, double plheat_frac) {
  d->thermal = on;
  d->k_T = k_T;
  d->cp_T = cp_T;
  d->plheat_frac = plheat_frac;
}

/* ThermalCalcs (Thermal.C:27-126): element conduction Kt(i,j) =
 * (sum_d dH(d,i) dH(d,j)) k_T / detJ * w on this step's Jacobians,
 * element-node rates -(Kt T)_i / m_i + 0.9 q_plheat vol / n / (m_i cp_T),
 * summed per node in CSR order, T += (dTdt + q_cont_conv / (m cp_T)) * dt
 * with the contact conduction of the last contact step (Thermal.C:124). */
void wfref_thermal_calcs(wfref_domain *d) {
  int n = d->nodxelem, E = d->elem_count, N = d->node_count;
  double w = (n == 4) ? 1.0 / 6.0 : 8.0;
#pragma omp parallel for
  for (int e = 0; e < E; ++e) {
    const double *hx = d->dHdx + n * e, *hy = d->dHdy + n * e, *hz = d->dHdz + n * e;
    double Te[8], dTde[8];
    for (int i = 0; i < n; ++i) Te[i] = d->T[d->elnod[n * e + i]];
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) {
        double k = 0.0;
        k += hx[i] * hx[j];
        k += hy[i] * hy[j];
        k += hz[i] * hz[j];
        double kt = k * d->k_T / d->detJ[e] * w;
        acc += kt * Te[j];
      }
      dTde[i] = acc;
    }
    double heat = 0.9 * d->q_plheat[e];
    double elem_pow = heat * d->vol[e];
    double pow_per_node = elem_pow / n;
    for (int i = 0; i < n; ++i) {
      int node = d->elnod[n * e + i];
      double m_inv = 1.0 / d->mdiag[node];
      double r = -m_inv * dTde[i];
      r += pow_per_node / (d->mdiag[node] * d->cp_T);
      d->dTedt[n * e + i] = r;
    }
  }
#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    double dTdt = 0.0;
    for (int k = 0; k < d->nodel_count[i]; ++k) {
      int o = d->nodel_offset[i] + k;
      dTdt += d->dTedt[d->nodel[o] * n + d->nodel_loc[o]];
    }
    d->T[i] += (dTdt + d->q_cont_conv[i] * 1.0 / (d->mdiag[i] * d->cp_T)) * d->dt;
  }
}

void wfref_set_thermal_expansion(wfref_domain *d, double exp_T) { d->exp_T = exp_T; }

/* calcThermalExpansion (Thermal.C:164-178): the element's mean of the
 * element-node temperature rates, taken as the reference writes it, at the
 * entries m_dTedt[node] of the element-node array (flat index = node id, the
 * rates the last ThermalCalcs left), is removed from the diagonal of the
 * strain rate: D = D - (exp_T dTdt) I. */
void wfref_calc_thermal_expansion(wfref_domain *d) {
  int n = d->nodxelem;
#pragma omp parallel for
  for (int e = 0; e < d->elem_count; ++e) {
    double dTdt_gp = 0.0;
    for (int i = 0; i < n; ++i) {
      int node = d->elnod[e * n + i];
      dTdt_gp += 1.0 / n * d->dTedt[node];
    }
    double s = d->exp_T * dTdt_gp;
    double *sr = d->str_rate + 6 * e;
    /* StrRate - s * Identity(): the diagonal loses s, the off-diagonal
     * entries s * 0 */
    sr[0] = sr[0] - s * 1.0;
    sr[1] = sr[1] - s * 1.0;
    sr[2] = sr[2] - s * 1.0;
    sr[3] = sr[3] - s * 0.0;
    sr[4] = sr[4] - s * 0.0;
    sr[5] = sr[5] - s * 0.0;
  }
}

/* Per-step time step of the BUILD_REMESH CPU solver (Solver_explicit.C:576-600):
 * after the prediction and the Jacobians, dt = cfl * min edge / (c + max |v|)
 * with c = sqrt(K / rho[0]) and |v| the predicted nodal speed; the step's
 * stress update and correction use it, the next step's prediction too. */
static void cfl_time_step(wfref_domain *d) {
  double mat_cs = sqrt(d->mat.K / d->rho[0]);
  double minl = wfref_calc_min_edge_length(d);
  double max_vel = 0.0;
  for (int i = 0; i < d->node_count; ++i) {
    const double *v = d->v + 3 * i;
    double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nv > max_vel) max_vel = nv;
  }
  d->dt = d->cfl * minl / (mat_cs + max_vel);
}


/* ------------------------------------------------------------------------ */
/* rigid-surface penalty contact                                            */
/* ------------------------------------------------------------------------ */
typedef struct { double x, y, z; } v3;
static v3 v3s(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 v3at(const double *p, int i) { return v3s(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static v3 v3sub(v3 a, v3 b) { return v3s(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 v3add(v3 a, v3 b) { return v3s(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 v3mul(v3 a, double f) { return v3s(a.x * f, a.y * f, a.z * f); }
static v3 v3smul(double f, v3 a) { return v3s(f * a.x, f * a.y, f * a.z); }
static v3 v3div(v3 a, double f) { return v3s(a.x / f, a.y / f, a.z / f); }
static double v3dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 v3cross(v3 a, v3 b) {
  return v3s(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static double v3len(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

/* CalcNormals (Mesh.h:236-248) and UpdatePlaneCoeff with nfar = 2
 * (CalcSpheres leaves the last node, Mesh.h:312-326; :275-283) */
static void rigid_planes(wfref_domain *d) {
  for (int e = 0; e < d->rt; ++e) {
    const int *t = d->rtri + 3 * e;
    v3 u = v3sub(v3at(d->rnode, t[1]), v3at(d->rnode, t[0]));
    v3 v = v3sub(v3at(d->rnode, t[2]), v3at(d->rnode, t[0]));
    v3 w = v3cross(u, v);
    v3 nrm = v3div(w, v3len(w));
    d->rnormal[3 * e] = nrm.x; d->rnormal[3 * e + 1] = nrm.y; d->rnormal[3 * e + 2] = nrm.z;
    d->rpplane[e] = v3dot(v3at(d->rnode, t[2]), nrm);
  }
}

/* tetra_faces (Domain_d.h:184-186) */
static const int TET_FACES[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

static int cmp_key(const void *a, const void *b) {
  const long long *x = (const long long *)a, *y = (const long long *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return 0;
}

/* SearchExtNodes (Domain_d.C:110-191): every tet face in element and
 * tetra_faces order, deduplicated by node set; faces met once are external,
 * kept in that order with the nodes of their element's table entry. */
static void search_ext_nodes(wfref_domain *d) {
  int E = d->elem_count, N = d->node_count, F = 4 * E;
  long long (*key)[2] = malloc(sizeof(*key) * (size_t)(F > 0 ? F : 1));
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 4; ++f) {
      int a = d->elnod[4 * e + TET_FACES[f][0]], b = d->elnod[4 * e + TET_FACES[f][1]],
          c = d->elnod[4 * e + TET_FACES[f][2]], t;
      if (a > b) { t = a; a = b; b = t; }
      if (b > c) { t = b; b = c; c = t; }
      if (a > b) { t = a; a = b; b = t; }
      key[4 * e + f][0] = ((long long)a * N + b) * (long long)N + c;
      key[4 * e + f][1] = 4 * e + f;
    }
  qsort(key, (size_t)F, sizeof(*key), cmp_key);
  int *cnt = calloc((size_t)(F > 0 ? F : 1), sizeof(int));
  for (int k = 0; k < F;) {
    int j = k;
    while (j < F && key[j][0] == key[k][0]) ++j;
    for (int q = k; q < j; ++q) cnt[key[q][1]] = j - k;
    k = j;
  }
  free(d->ext_face);
  d->ext_face = malloc(sizeof(int) * 3 * (size_t)(F > 0 ? F : 1));
  d->nf = 0;
  for (int i = 0; i < N; ++i) d->ext_node[i] = 0;
  for (int g = 0; g < F; ++g) {
    if (cnt[g] != 1) continue;
    int e = g / 4, f = g % 4;
    for (int k = 0; k < 3; ++k) {
      int nd = d->elnod[4 * e + TET_FACES[f][k]];
      d->ext_face[3 * d->nf + k] = nd;
      d->ext_node[nd] = 1;
    }
    d->nf++;
  }
  free(cnt);
  free(key);
}

/* CalcExtFaceAreas (Domain_d.C:207-296), 3D: a third of each external
 * triangle's area to each of its nodes */
void wfref_calc_ext_face_areas(wfref_domain *d) {
  for (int i = 0; i < d->node_count; ++i) d->node_area[i] = 0.0;
  for (int f = 0; f < d->nf; ++f) {
    const int *t = d->ext_face + 3 * f;
    v3 p0 = v3at(d->x, t[0]), p1 = v3at(d->x, t[1]), p2 = v3at(d->x, t[2]);
    v3 c = v3cross(v3sub(p1, p0), v3sub(p2, p0));
    double area = 0.5 * v3len(c);
    double share = area / 3.0;
    d->node_area[t[0]] += share;
    d->node_area[t[1]] += share;
    d->node_area[t[2]] += share;
  }
}

/* m_elem_length of calcMinEdgeLength (Domain_d.C:2242-2276), tetrahedra:
 * the smallest height over the four faces */
void wfref_calc_elem_length(wfref_domain *d) {
  for (int e = 0; e < d->elem_count; ++e) {
    const int *en = d->elnod + 4 * e;
    v3 P[4];
    for (int k = 0; k < 4; ++k) P[k] = v3at(d->x, en[k]);
    const int fc[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    double hmin = 1.0e6;
    for (int i = 0; i < 4; ++i) {
      v3 x1 = v3sub(P[fc[i][1]], P[fc[i][0]]), x2 = v3sub(P[fc[i][2]], P[fc[i][0]]);
      v3 nrm = v3cross(x1, x2);
      double area = v3len(nrm);
      if (area < 1e-12) continue;
      nrm = v3div(nrm, area);
      double h = fabs(v3dot(v3sub(P[i], P[fc[i][0]]), nrm));
      if (h < hmin) hmin = h;
    }
    d->elem_length[e] = hmin;
  }
}

int wfref_set_rigid_mesh(wfref_domain *d, int nn, const double *nodes, const double *vel, int nt,
                         const int *tris, const int *mesh_id) {
  if (d->nodxelem != 4 || nn <= 0 || nt <= 0) return -1;
  free_rigid(d);
  d->rn = nn; d->rt = nt;
  RD(d->rnode, 3 * nn); RD(d->rvel, 3 * nn); RI(d->rtri, 3 * nt); RI(d->rmesh, nt);
  RD(d->rnormal, 3 * nt); RD(d->rpplane, nt);
  memcpy(d->rnode, nodes, sizeof(double) * 3 * (size_t)nn);
  if (vel) memcpy(d->rvel, vel, sizeof(double) * 3 * (size_t)nn);
  memcpy(d->rtri, tris, sizeof(int) * 3 * (size_t)nt);
  if (mesh_id) memcpy(d->rmesh, mesh_id, sizeof(int) * (size_t)nt);
  rigid_planes(d);
  search_ext_nodes(d);
  wfref_calc_ext_face_areas(d);
  d->contact = 1;
  return 0;
}

/* TriMesh_d::Move + CalcNormals + UpdatePlaneCoeff after the position
 * update (Solver_explicit.C:996-1004) */
void wfref_rigid_update(wfref_domain *d) {
  for (int n = 0; n < d->rn; ++n)
    for (int c = 0; c < 3; ++c) d->rnode[3 * n + c] = d->rnode[3 * n + c] + d->rvel[3 * n + c] * d->dt;
  rigid_planes(d);
}

/* CalcContactForces (Contact.C:31-335): for every external node, the first
 * rigid triangle (in order) it lies behind and inside of gives a penalty
 * force along the triangle normal with critical damping ratio 0.2, plus a
 * stick/slip Coulomb friction force from the accumulated tangential
 * displacement.  Nodes out of contact keep their previous contact force
 * (the reference's reset is commented out, Contact.C:54). */
void wfref_calc_contact_forces(wfref_domain *d) {
  int N = d->node_count;
  double dt = d->dt;
#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    d->mesh_in_contact[i] = -1;
    if (!d->ext_node[i]) continue;
    v3 xi = v3at(d->x, i), vi = v3at(d->v, i);
    for (int j = 0; j < d->rt; ++j) {
      v3 nj = v3at(d->rnormal, j);
      double dd = v3dot(nj, xi) - d->rpplane[j];
      if (!(dd < 0)) continue;
      v3 Qj = v3sub(xi, v3smul(dd, nj));
      int inside = 1;
      for (int l = 0; l < 3 && inside; ++l) {
        int m = l + 1;
        if (m > 2) m = 0;
        v3 Pl = v3at(d->rnode, d->rtri[3 * j + l]), Pm = v3at(d->rnode, d->rtri[3 * j + m]);
        double crit = v3dot(v3cross(v3sub(Pm, Pl), v3sub(Qj, Pl)), nj);
        if (crit < 0.0) inside = 0;
      }
      if (!inside) continue;
      double nodlen = 0.0;
      for (int k = 0; k < d->nodel_count[i]; ++k)
        nodlen += d->elem_length[d->nodel[d->nodel_offset[i] + k]];
      nodlen /= d->nodel_count[i];
      double m = d->mdiag[i];
      double v_reln = v3dot(vi, nj);
      double kcont_geo = d->mat.E * d->node_area[i] / nodlen;
      double kcont_mass = 0.2 * m / (dt * dt);
      double kcont = (kcont_mass < kcont_geo) ? kcont_mass : kcont_geo; /* std::min */
      double omega = sqrt(kcont / m);
      double ccrit = 2.0 * m * omega;
      double F_damp = 0.2 * ccrit * v_reln;
      double F_normal = d->contPF * kcont * dd;
      v3 cf = v3smul(-(F_normal + F_damp), nj);
      double *cfo = d->contforce + 3 * i;
      cfo[0] = cf.x; cfo[1] = cf.y; cfo[2] = cf.z;
      d->cfn[i] = v3len(cf);
      d->mesh_in_contact[i] = d->rmesh[j];
      /* friction (Contact.C:248-306) */
      v3 v_tan = v3sub(vi, v3smul(v3dot(vi, nj), nj));
      v3 du_t = v3mul(v_tan, dt);
      double *ut = d->ut_prev + 3 * i;
      v3 ut_acc = v3add(v3s(ut[0], ut[1], ut[2]), du_t);
      ut[0] += du_t.x; ut[1] += du_t.y; ut[2] += du_t.z;
      v3 Ft_trial = v3smul(-kcont, ut_acc);
      double Ft_mag = v3len(Ft_trial);
      v3 Fn = v3smul(v3dot(cf, nj), nj);
      double Ft_max_static = d->mu_sta * v3len(Fn);
      v3 Ft;
      if (Ft_mag <= Ft_max_static) {
        Ft = Ft_trial;
      } else {
        double Ft_max_dynamic = d->mu_dyn * v3len(Fn);
        Ft = v3div(v3smul(-Ft_max_dynamic, v_tan), v3len(v_tan));
        ut[0] = ut[1] = ut[2] = 0.0;
      }
      cfo[0] += Ft.x; cfo[1] += Ft.y; cfo[2] += Ft.z;
      /* contact heat conduction (Contact.C:309) */
      d->q_cont_conv[i] = d->heat_cond * d->node_area[i] * (d->T_const - d->T[i]);
      break; /* one master element per slave node */
    }
  }
}

/* trimesh->react_force[m].z of Contact.C:318-321: sum of |normal + damping
 * force| over the nodes in contact with rigid mesh m, in node order */
double wfref_contact_force(wfref_domain *d, int mesh_id) {
  double r = 0.0;
  for (int i = 0; i < d->node_count; ++i)
    if (d->mesh_in_contact[i] == mesh_id) r += d->cfn[i];
  return r;
}

/* One pass of the SolveChungHulbert loop body, CPU branch, in its order
 * (Solver_explicit.C:524-1012; remesh and I/O left out). */
void wfref_step(wfref_domain *d, double *reaction_z) {
  /* external face areas every 10 steps (Solver_explicit.C:446-451) */
  if (d->contact && d->step_count % 10 == 0) wfref_calc_ext_face_areas(d);
  wfref_update_prediction(d);
  for (int c = 0; c < 3; ++c) wfref_impose_bcv(d, c);
  wfref_calc_elem_jacobians(d);
  if (d->cfl > 0.0) {
    cfl_time_step(d);
    if (d->contact) wfref_calc_elem_length(d); /* calcMinEdgeLength */
  }
  wfref_calc_elem_vol(d);
  wfref_calc_nodal_vol(d);
  wfref_calc_nodal_mass_from_vol(d);
  wfref_calc_elem_strain_rates(d);
  if (d->thermal) wfref_calc_thermal_expansion(d); /* Solver_explicit.C:719-720 */
  wfref_calc_elem_pressure(d);
  wfref_calc_stress_strain(d, d->dt);
  wfref_calc_artificial_viscosity(d);
  wfref_calc_elem_forces(d);
  wfref_calc_elem_hourglass_forces(d);
  if (d->contact) wfref_calc_contact_forces(d);
  wfref_assembly_forces(d);
  wfref_scrub_nonfinite_fi(d);
  if (reaction_z) *reaction_z = wfref_reaction_force(d, 2);
  wfref_calc_accel(d);
  for (int c = 0; c < 3; ++c) wfref_impose_bca(d, c);
  wfref_update_correction_accvel(d);
  for (int c = 0; c < 3; ++c) wfref_impose_bcv(d, c);
  wfref_update_correction_pos(d);
  if (d->contact) wfref_rigid_update(d);
  if (d->thermal) wfref_thermal_calcs(d);
  d->step_count++;
}

void wfref_run(wfref_domain *d, int nsteps, double *hist) {
  for (int s = 0; s < nsteps; ++s) wfref_step(d, hist ? hist + s : NULL);
}

/* computeEnergies (Mechanical.C:2145-2185) */
void wfref_energies(wfref_domain *d, double *ekin, double *deint) {
  double ek = 0.0;
  for (int i = 0; i < d->node_count; ++i) {
    double vx = d->v[3 * i], vy = d->v[3 * i + 1], vz = d->v[3 * i + 2];
    ek += 0.5 * d->mdiag[i] * (vx * vx + vy * vy + vz * vz);
  }
  double ed = 0.0;
  for (int e = 0; e < d->elem_count; ++e) {
    const double *s = d->sigma + 6 * e, *r = d->str_rate + 6 * e;
    double sd = s[0] * r[0] + s[1] * r[1] + s[2] * r[2] + s[3] * r[3] + s[4] * r[4] + s[5] * r[5];
    ed += sd * d->vol[e];
  }
  *ekin = ek;
  *deint = ed * d->dt;
}

/* field access by name for the Python test harness (reference layout) */
void *wfref_field(wfref_domain *d, const char *name, long *count, int *is_int) {
  long N = d->node_count, E = d->elem_count, n = d->nodxelem;
  struct { const char *nm; void *p; long c; int i; } t[] = {
      {"x", d->x, 3 * N, 0}, {"v", d->v, 3 * N, 0}, {"a", d->a, 3 * N, 0},
      {"u", d->u, 3 * N, 0}, {"u_dt", d->u_dt, 3 * N, 0}, {"prev_a", d->prev_a, 3 * N, 0},
      {"fi", d->fi, 3 * N, 0}, {"fe", d->fe, 3 * N, 0}, {"mdiag", d->mdiag, N, 0},
      {"voln", d->voln, N, 0}, {"elnod", d->elnod, n * E, 1}, {"nodel", d->nodel, n * E, 1},
      {"nodel_loc", d->nodel_loc, n * E, 1}, {"nodel_offset", d->nodel_offset, N, 1},
      {"nodel_count", d->nodel_count, N, 1}, {"detJ", d->detJ, E, 0},
      {"dHdx", d->dHdx, n * E, 0}, {"dHdy", d->dHdy, n * E, 0}, {"dHdz", d->dHdz, n * E, 0},
      {"vol", d->vol, E, 0}, {"vol0", d->vol0, E, 0}, {"rho", d->rho, E, 0},
      {"rho0", d->rho0, E, 0}, {"p", d->p, E, 0}, {"pl_strain", d->pl_strain, E, 0},
      {"sigma_y", d->sigma_y, E, 0}, {"sy0", d->sy0, E, 0}, {"temp", d->temp, E, 0}, {"T", d->T, N, 0},
      {"q_plheat", d->q_plheat, E, 0}, {"dTedt", d->dTedt, n * E, 0},
      {"node_area", d->node_area, N, 0}, {"elem_length", d->elem_length, E, 0},
      {"contforce", d->contforce, 3 * N, 0}, {"ut_prev", d->ut_prev, 3 * N, 0},
      {"cfn", d->cfn, N, 0}, {"mesh_in_contact", d->mesh_in_contact, N, 1},
      {"q_cont_conv", d->q_cont_conv, N, 0},
      {"ext_node", d->ext_node, N, 1}, {"rigid_x", d->rnode, 3 * (long)d->rn, 0},
      {"rigid_v", d->rvel, 3 * (long)d->rn, 0}, {"rigid_normal", d->rnormal, 3 * (long)d->rt, 0},
      {"rigid_pplane", d->rpplane, d->rt, 0},
      {"str_rate", d->str_rate, 6 * E, 0}, {"rot_rate", d->rot_rate, 6 * E, 0},
      {"sigma", d->sigma, 6 * E, 0}, {"tau", d->tau, 6 * E, 0}, {"eps", d->eps, 6 * E, 0},
      {"strain_pl_incr", d->strain_pl_incr, 6 * E, 0}, {"f_elem", d->f_elem, 3 * n * E, 0},
      {"f_hg", d->f_hg, 3 * n * E, 0}};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); ++k)
    if (strcmp(t[k].nm, name) == 0) {
      *count = t[k].c;
      *is_int = t[k].i;
      return t[k].p;
    }
  *count = 0;
  *is_int = 0;
  return NULL;
}

int wfref_get_int(wfref_domain *d, const char *name) {
  if (!strcmp(name, "node_count")) return d->node_count;
  if (!strcmp(name, "elem_count")) return d->elem_count;
  if (!strcmp(name, "nodxelem")) return d->nodxelem;
  if (!strcmp(name, "nonfinite_fi")) return d->nonfinite_fi;
  if (!strcmp(name, "inverted")) return d->inverted;
  return -1;
}

double wfref_get_double(wfref_domain *d, const char *name) {
  if (!strcmp(name, "dt")) return d->dt;
  if (!strcmp(name, "alpha")) return d->alpha;
  if (!strcmp(name, "beta")) return d->beta;
  if (!strcmp(name, "gamma")) return d->gamma;
  if (!strcmp(name, "K")) return d->mat.K;
  if (!strcmp(name, "G")) return d->mat.G;
  if (!strcmp(name, "cs0")) return d->mat.cs0;
  if (!strcmp(name, "H")) return d->mat.H;
  if (!strcmp(name, "max_edot")) return d->max_edot;
  return NAN;
}

int wfref_set_double(wfref_domain *d, const char *name, double v) {
  struct { const char *nm; double *p; } t[] = {
      {"dt", &d->dt}, {"alpha_free", &d->alpha_free}, {"log_factor", &d->log_factor},
      {"J_min", &d->J_min}, {"hg_coeff_free", &d->hg_coeff_free},
      {"alpha_contact", &d->alpha_contact}, {"hg_coeff_contact", &d->hg_coeff_contact},
      {"pspg_scale", &d->pspg_scale}, {"p_pspg_bulkfac", &d->p_pspg_bulkfac},
      {"av_coeff_div", &d->av_coeff_div}, {"av_coeff_bulk", &d->av_coeff_bulk},
      {"artifvisc0", &d->artifvisc[0]}, {"artifvisc1", &d->artifvisc[1]},
      {"hg_coeff", &d->hg_coeff}, {"max_edot", &d->max_edot}, {"cfl", &d->cfl},
      {"contPF", &d->contPF}, {"mu_sta", &d->mu_sta}, {"mu_dyn", &d->mu_dyn},
      {"heat_cond", &d->heat_cond}, {"T_const", &d->T_const}};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); ++k)
    if (strcmp(t[k].nm, name) == 0) { *t[k].p = v; return 0; }
  return -1;
}

#ifdef _OPENMP
#include <omp.h>
#endif
/* OpenMP threads the oracle's parallel loops use */
int wfref_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
