// FP64 instantiation of the per-step kernels (built with -ffp-contract=off)
// and the precision-independent nodal launchers; the fused FP64 step is
// instantiated in wfx_f64.hip.
#include "wfx_kernels.hpp"
#include "wfx_fused.hpp"
#include "wfx_launch.hpp"

namespace wfx {

WFX_DEFINE_LAUNCHERS(double)
template hipError_t launch_cfl_step<double>(wfx_ctx *, int, hipStream_t);
template hipError_t launch_thermal<double>(wfx_ctx *, hipStream_t);
template hipError_t launch_thermal_expansion<double>(wfx_ctx *, hipStream_t);

hipError_t launch_update_prediction(wfx_ctx *c, hipStream_t s) {
  long n3 = 3L * c->N;
  const auto &p = c->prm;
  // with the per-step CFL time step, dt is the device's value of the last step
  const double *dtp = p.cfl > 0.0 ? c->scalars + kCflDt : nullptr;
  hipLaunchKernelGGL(k_update_prediction, dim3(nblocks(n3)), dim3(kBlock), 0, s, n3, c->nd.v,
                     c->nd.prev_a, c->nd.u_dt, p.dt, p.beta, p.gamma, dtp);
  return hipGetLastError();
}

namespace {
// plain device copy as a kernel (usable inside captured graphs)
__global__ __launch_bounds__(kBlock) void k_copy(long n, const double *__restrict__ src,
                                                 double *__restrict__ dst) {
  long i = gtid();
  if (i < n) dst[i] = src[i];
}
}  // namespace

hipError_t launch_copy(const double *src, double *dst, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, s, n, src, dst);
  return hipGetLastError();
}

hipError_t launch_impose(wfx_ctx *c, int dim, bool accel, hipStream_t s) {
  int cnt = c->bc_count[dim];
  if (cnt == 0) return hipSuccess;
  double *comp = (accel ? c->nd.a : c->nd.v) + (long)dim * c->N;
  hipLaunchKernelGGL(k_impose_value, dim3(nblocks(cnt)), dim3(kBlock), 0, s, cnt, c->bc_nod[dim],
                     accel ? (const double *)nullptr : (const double *)c->bc_val[dim], comp);
  return hipGetLastError();
}

hipError_t launch_calc_accel(wfx_ctx *c, hipStream_t s) {
  hipLaunchKernelGGL(k_calc_accel, dim3(nblocks(c->N)), dim3(kBlock), 0, s, c->N, c->nd.fe, c->nd.fi,
                     c->nd.mdiag, c->nd.a, c->contact.on ? c->contact.contforce : nullptr);
  return hipGetLastError();
}

hipError_t launch_correction_accvel(wfx_ctx *c, hipStream_t s) {
  long n3 = 3L * c->N;
  const auto &p = c->prm;
  double f = 1.0 / (1.0 - p.alpha);
  hipLaunchKernelGGL(k_correction_accvel, dim3(nblocks(n3)), dim3(kBlock), 0, s, n3, c->nd.a,
                     c->nd.prev_a, c->nd.v, f, p.alpha, p.gamma * p.dt);
  return hipGetLastError();
}

hipError_t launch_correction_pos(wfx_ctx *c, hipStream_t s) {
  long n3 = 3L * c->N;
  const auto &p = c->prm;
  hipLaunchKernelGGL(k_correction_pos, dim3(nblocks(n3)), dim3(kBlock), 0, s, n3, c->nd.a,
                     c->nd.u_dt, c->nd.x, c->nd.prev_a, c->nd.u, p.beta * p.dt * p.dt);
  return hipGetLastError();
}

hipError_t launch_reaction(wfx_ctx *c, int dim, double *dev_out, hipStream_t s) {
  if (c->ngroups == 0) return hipSuccess;
  hipLaunchKernelGGL((k_reaction<double>), dim3(c->ngroups), dim3(kBlock), 0, s, c->react_off, c->react,
                     c->nd.fi + (long)dim * c->N, dev_out);
  return hipGetLastError();
}

hipError_t launch_reaction_f32(wfx_ctx *c, const float *fi32, double *dev_out, hipStream_t s) {
  if (c->ngroups == 0) return hipSuccess;
  hipLaunchKernelGGL((k_reaction<float>), dim3(c->ngroups), dim3(kBlock), 0, s, c->react_off, c->react,
                     fi32 + 2L * c->N, dev_out);
  return hipGetLastError();
}

int min_edge_partial_count(wfx_ctx *c) { return nblocks(c->E); }

hipError_t launch_min_edge(wfx_ctx *c, double *partial, double *dev_out, hipStream_t s) {
  int nb = nblocks(c->E);
  hipLaunchKernelGGL(k_min_edge, dim3(nb), dim3(kBlock), 0, s, c->E, c->N, c->n, c->ms.elnod, c->nd.x,
                     partial);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kBlock), 0, s, nb, partial, dev_out);
  return hipGetLastError();
}

// contact kernels for the per-step API (the fused path launches them itself)
hipError_t contact_face_areas(wfx_ctx *c, hipStream_t s) { return launch_ext_face_areas(c, s); }
hipError_t contact_elem_length(wfx_ctx *c, hipStream_t s) { return launch_elem_length(c, s); }
hipError_t contact_forces(wfx_ctx *c, hipStream_t s) { return launch_contact<double>(c, c->nd.v, false, s); }
hipError_t contact_rigid_update(wfx_ctx *c, double dt, hipStream_t s) {
  const double keep = c->prm.dt;
  c->prm.dt = dt;
  hipError_t e = launch_rigid_update(c, false, s);
  c->prm.dt = keep;
  return e;
}

}  // namespace wfx
