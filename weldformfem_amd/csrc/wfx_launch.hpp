// Declarations of the kernel launchers.  The precision-dependent ones are
// defined (explicit specialisations) in wfx_k64.hip (double) and wfx_k32.hip
// (float); the nodal ones, which are precision independent, in wfx_k64.hip.
#pragma once

#include "wfx_internal.hpp"

namespace wfx {

template <typename T> hipError_t launch_elem_jacobians(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_elem_vol(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_elem_density(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_nodal_vol(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_nodal_mass(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_strain_rates(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_pressure(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_stress_strain(wfx_ctx *c, double dt, hipStream_t s);
template <typename T> hipError_t launch_artificial_viscosity(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_elem_forces(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_hourglass(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_assembly(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_energies(wfx_ctx *c, double *dev_out2, double *partial, hipStream_t s);

#define WFX_DECLARE_SPECIALISATIONS(T)                                                          \
  template <> hipError_t launch_elem_jacobians<T>(wfx_ctx *, hipStream_t);                       \
  template <> hipError_t launch_elem_vol<T>(wfx_ctx *, hipStream_t);                             \
  template <> hipError_t launch_elem_density<T>(wfx_ctx *, hipStream_t);                         \
  template <> hipError_t launch_nodal_vol<T>(wfx_ctx *, hipStream_t);                            \
  template <> hipError_t launch_nodal_mass<T>(wfx_ctx *, hipStream_t);                           \
  template <> hipError_t launch_strain_rates<T>(wfx_ctx *, hipStream_t);                         \
  template <> hipError_t launch_pressure<T>(wfx_ctx *, hipStream_t);                             \
  template <> hipError_t launch_stress_strain<T>(wfx_ctx *, double, hipStream_t);                \
  template <> hipError_t launch_artificial_viscosity<T>(wfx_ctx *, hipStream_t);                 \
  template <> hipError_t launch_elem_forces<T>(wfx_ctx *, hipStream_t);                          \
  template <> hipError_t launch_hourglass<T>(wfx_ctx *, hipStream_t);                            \
  template <> hipError_t launch_assembly<T>(wfx_ctx *, hipStream_t);                             \
  template <> hipError_t launch_energies<T>(wfx_ctx *, double *, double *, hipStream_t);
WFX_DECLARE_SPECIALISATIONS(double)
WFX_DECLARE_SPECIALISATIONS(float)

// ev (optional, 4 * nsteps + 1 events): recorded before each phase of each
// step and after the last, for per-phase durations of a run
template <typename T>
hipError_t launch_fused_steps(wfx_ctx *c, int nsteps, double *hist, hipStream_t s, hipEvent_t *ev);
template <typename T> hipError_t launch_vsum0(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_fused_phase(wfx_ctx *c, int phase, int predict, hipStream_t s, int step);
template <typename T> hipError_t launch_cfl_step(wfx_ctx *c, int step, hipStream_t s);
template <typename T> hipError_t launch_thermal(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_thermal_expansion(wfx_ctx *c, hipStream_t s);
extern template hipError_t launch_thermal_expansion<double>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_thermal_expansion<float>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_thermal<double>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_thermal<float>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_cfl_step<double>(wfx_ctx *, int, hipStream_t);
extern template hipError_t launch_cfl_step<float>(wfx_ctx *, int, hipStream_t);
extern template hipError_t launch_fused_steps<double>(wfx_ctx *, int, double *, hipStream_t, hipEvent_t *);
extern template hipError_t launch_fused_steps<float>(wfx_ctx *, int, double *, hipStream_t, hipEvent_t *);
extern template hipError_t launch_vsum0<double>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_vsum0<float>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_fused_phase<double>(wfx_ctx *, int, int, hipStream_t, int);
extern template hipError_t launch_fused_phase<float>(wfx_ctx *, int, int, hipStream_t, int);
template <typename T> hipError_t launch_fused_pack(wfx_ctx *c, hipStream_t s);
template <typename T> hipError_t launch_cfl_begin(wfx_ctx *c, hipStream_t s, bool packed);
extern template hipError_t launch_cfl_begin<double>(wfx_ctx *, hipStream_t, bool);
extern template hipError_t launch_cfl_begin<float>(wfx_ctx *, hipStream_t, bool);
template <typename T> hipError_t launch_fused_unpack(wfx_ctx *c, hipStream_t s);
extern template hipError_t launch_fused_pack<double>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_fused_pack<float>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_fused_unpack<double>(wfx_ctx *, hipStream_t);
extern template hipError_t launch_fused_unpack<float>(wfx_ctx *, hipStream_t);

hipError_t launch_update_prediction(wfx_ctx *c, hipStream_t s);
hipError_t launch_copy(const double *src, double *dst, long n, hipStream_t s);
hipError_t launch_impose(wfx_ctx *c, int dim, bool accel, hipStream_t s);
hipError_t launch_calc_accel(wfx_ctx *c, hipStream_t s);
hipError_t launch_correction_accvel(wfx_ctx *c, hipStream_t s);
hipError_t launch_correction_pos(wfx_ctx *c, hipStream_t s);
hipError_t launch_reaction(wfx_ctx *c, int dim, double *dev_out, hipStream_t s);
hipError_t launch_min_edge(wfx_ctx *c, double *partial, double *dev_out, hipStream_t s);
int min_edge_partial_count(wfx_ctx *c);
hipError_t launch_reaction_f32(wfx_ctx *c, const float *fi32, double *dev_out, hipStream_t s);

// rigid-surface contact, per-step API (wfx_k64.hip)
hipError_t contact_face_areas(wfx_ctx *c, hipStream_t s);
hipError_t contact_elem_length(wfx_ctx *c, hipStream_t s);
hipError_t contact_forces(wfx_ctx *c, hipStream_t s);
hipError_t contact_rigid_update(wfx_ctx *c, double dt, hipStream_t s);

// device buffers a fused run writes (saved and restored around autotuning
// and around the adjoint's recomputation)
std::vector<std::pair<void *, size_t>> fused_state_buffers(wfx_ctx *c);
void adjoint_free(wfx_ctx *c);
void frames_free(wfx_ctx *c);
void frames_join(wfx_ctx *c);

}  // namespace wfx
