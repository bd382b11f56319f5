// Rigid-surface penalty contact (Contact.C:31-335) for a tetrahedral body
// against rigid triangle surfaces (TriMesh_d, Mesh.h): nodal areas of the
// external faces (CalcExtFaceAreas, Domain_d.C:207-296), element heights
// (m_elem_length of calcMinEdgeLength, Domain_d.C:2242-2276), the contact
// force of every external node (brute-force search over the rigid
// triangles, first hit wins) and the rigid surfaces' motion (Move,
// CalcNormals, UpdatePlaneCoeff, Mesh.h:236-330).  Included by
// wfx_fused.hpp; every expression follows the reference's operation order.
#pragma once

#include "wfx_kernels.hpp"

namespace wfx {
namespace {

struct R3 {
  double x, y, z;
};
__device__ __forceinline__ R3 r3(double x, double y, double z) { return R3{x, y, z}; }
__device__ __forceinline__ R3 r3_aos(const double *p, long i) { return R3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ R3 r3_soa(const double *p, long N, long i) {
  return R3{p[i], p[N + i], p[2 * N + i]};
}
__device__ __forceinline__ R3 operator-(R3 a, R3 b) { return r3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ R3 operator+(R3 a, R3 b) { return r3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ R3 operator*(double f, R3 a) { return r3(f * a.x, f * a.y, f * a.z); }
__device__ __forceinline__ R3 operator*(R3 a, double f) { return r3(a.x * f, a.y * f, a.z * f); }
__device__ __forceinline__ R3 operator/(R3 a, double f) { return r3(a.x / f, a.y / f, a.z / f); }
__device__ __forceinline__ double dot(R3 a, R3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ R3 cross(R3 a, R3 b) {
  return r3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double len(R3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// node_area of every external node: a third of each of its external
// triangles' areas, summed in face order
__global__ __launch_bounds__(kBlock) void k_ext_face_areas(int next, const int *__restrict__ ext_nodes,
                                                           const int *__restrict__ afc_off,
                                                           const int *__restrict__ afc_face,
                                                           const int *__restrict__ ext_face, int N,
                                                           const double *__restrict__ x,
                                                           double *__restrict__ node_area) {
  const int k = gtid();
  if (k >= next) return;
  double s = 0.0;
  for (int q = afc_off[k]; q < afc_off[k + 1]; ++q) {
    const int *t = ext_face + 3 * afc_face[q];
    const R3 p0 = r3_soa(x, N, t[0]), p1 = r3_soa(x, N, t[1]), p2 = r3_soa(x, N, t[2]);
    const double area = 0.5 * len(cross(p1 - p0, p2 - p0));
    const double share = area / 3.0;
    s += share;
  }
  node_area[ext_nodes[k]] = s;
}

// smallest height of every tetrahedron
__global__ __launch_bounds__(kBlock) void k_elem_length(int E, int N, const int *__restrict__ elnod,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ elem_length) {
  const int e = gtid();
  if (e >= E) return;
  R3 P[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) P[k] = r3_soa(x, N, elnod[(long)k * E + e]);
  const int fc[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  double hmin = 1.0e6;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R3 nrm = cross(P[fc[i][1]] - P[fc[i][0]], P[fc[i][2]] - P[fc[i][0]]);
    const double area = len(nrm);
    if (area < 1e-12) continue;
    nrm = nrm / area;
    const double h = fabs(dot(P[i] - P[fc[i][0]], nrm));
    if (h < hmin) hmin = h;
  }
  elem_length[e] = hmin;
}

struct ContactParams {
  int rt;
  const double *rnode, *rnormal, *rpplane;
  const int *rtri, *rmesh;
  double E, pf, mu_sta, mu_dyn, dt;
  const double *dtp;  // per-step CFL time step on the device, or null
  // contact heat conduction (with thermal coupling; q_cont null: none)
  double heat_cond, T_const;
  const double *Tn;
  double *q_cont;
};

// CalcContactForces for every external node (v: the predicted velocities,
// stored in T); nodes out of contact keep their contact force and their
// contact heat conduction
template <typename T>
__global__ __launch_bounds__(kBlock) void k_contact(
    int next, const int *__restrict__ ext_nodes, int N, ContactParams cp, const double *__restrict__ x,
    const T *__restrict__ v, const double *__restrict__ mdiag, const int *__restrict__ off,
    const int *__restrict__ cnt, const int *__restrict__ nodel, const double *__restrict__ elem_length,
    const double *__restrict__ node_area, double *__restrict__ contforce, double *__restrict__ ut_prev,
    int *__restrict__ mesh_in_contact, double *__restrict__ cfn) {
  const int k = gtid();
  if (k >= next) return;
  const int i = ext_nodes[k];
  const double dt = cp.dtp ? *cp.dtp : cp.dt;
  mesh_in_contact[i] = -1;
  const R3 xi = r3_soa(x, N, i);
  const R3 vi = r3((double)v[i], (double)v[N + i], (double)v[2L * N + i]);
  for (int j = 0; j < cp.rt; ++j) {
    const R3 nj = r3_aos(cp.rnormal, j);
    const double d = dot(nj, xi) - cp.rpplane[j];
    if (!(d < 0)) continue;
    const R3 Qj = xi - d * nj;
    bool inside = true;
    for (int l = 0; l < 3 && inside; ++l) {
      const int m = (l + 1 > 2) ? 0 : l + 1;
      const R3 Pl = r3_aos(cp.rnode, cp.rtri[3 * j + l]), Pm = r3_aos(cp.rnode, cp.rtri[3 * j + m]);
      if (dot(cross(Pm - Pl, Qj - Pl), nj) < 0.0) inside = false;
    }
    if (!inside) continue;
    double nodlen = 0.0;
    const int o = off[i], c = cnt[i];
    for (int q = 0; q < c; ++q) nodlen += elem_length[nodel[o + q]];
    nodlen /= c;
    const double m = mdiag[i];
    const double v_reln = dot(vi, nj);
    const double kcont_geo = cp.E * node_area[i] / nodlen;
    const double kcont_mass = 0.2 * m / (dt * dt);
    const double kcont = (kcont_mass < kcont_geo) ? kcont_mass : kcont_geo;  // std::min
    const double omega = sqrt(kcont / m);
    const double ccrit = 2.0 * m * omega;
    const double F_damp = 0.2 * ccrit * v_reln;
    const double F_normal = cp.pf * kcont * d;
    const R3 cf = (-(F_normal + F_damp)) * nj;
    cfn[i] = len(cf);
    mesh_in_contact[i] = cp.rmesh[j];
    // friction (Contact.C:248-306)
    const R3 v_tan = vi - dot(vi, nj) * nj;
    const R3 du_t = v_tan * dt;
    R3 ut = r3_soa(ut_prev, N, i);
    const R3 ut_acc = ut + du_t;
    ut = r3(ut.x + du_t.x, ut.y + du_t.y, ut.z + du_t.z);
    const R3 Ft_trial = (-kcont) * ut_acc;
    const double Ft_mag = len(Ft_trial);
    const R3 Fn = dot(cf, nj) * nj;
    const double Ft_max_static = cp.mu_sta * len(Fn);
    R3 Ft;
    if (Ft_mag <= Ft_max_static) {
      Ft = Ft_trial;
    } else {
      const double Ft_max_dynamic = cp.mu_dyn * len(Fn);
      Ft = ((-Ft_max_dynamic) * v_tan) / len(v_tan);
      ut = r3(0.0, 0.0, 0.0);
    }
    ut_prev[i] = ut.x;
    ut_prev[N + i] = ut.y;
    ut_prev[2L * N + i] = ut.z;
    contforce[i] = cf.x + Ft.x;
    contforce[N + i] = cf.y + Ft.y;
    contforce[2L * N + i] = cf.z + Ft.z;
    // contact heat conduction (Contact.C:309), kept by a node that leaves contact
    if (cp.q_cont) cp.q_cont[i] = cp.heat_cond * node_area[i] * (cp.T_const - cp.Tn[i]);
    break;  // one master element per slave node
  }
}

// rigid surfaces after the position update: node += v dt, then normals and
// plane offsets (nfar = 2); one block
__global__ __launch_bounds__(kBlock) void k_rigid_update(int rn, int rt, double *__restrict__ rnode,
                                                         const double *__restrict__ rvel,
                                                         const int *__restrict__ rtri,
                                                         double *__restrict__ rnormal,
                                                         double *__restrict__ rpplane, double dt,
                                                         const double *__restrict__ dtp) {
  if (dtp) dt = *dtp;
  for (int k = threadIdx.x; k < 3 * rn; k += blockDim.x) rnode[k] = rnode[k] + rvel[k] * dt;
  __syncthreads();
  for (int e = threadIdx.x; e < rt; e += blockDim.x) {
    const int *t = rtri + 3 * e;
    const R3 w = cross(r3_aos(rnode, t[1]) - r3_aos(rnode, t[0]), r3_aos(rnode, t[2]) - r3_aos(rnode, t[0]));
    const R3 nrm = w / len(w);
    rnormal[3 * e] = nrm.x;
    rnormal[3 * e + 1] = nrm.y;
    rnormal[3 * e + 2] = nrm.z;
    rpplane[e] = dot(r3_aos(rnode, t[2]), nrm);
  }
}

}  // namespace
}  // namespace wfx
