// FP32-mixed instantiation of the per-step element kernels.
#include "wfx_kernels.hpp"
#include "wfx_fused.hpp"
#include "wfx_launch.hpp"

namespace wfx {
WFX_DEFINE_LAUNCHERS(float)
template hipError_t launch_vsum0<float>(wfx_ctx *, hipStream_t);
template hipError_t launch_fused_phase<float>(wfx_ctx *, int, int, hipStream_t, int);
template hipError_t launch_cfl_step<float>(wfx_ctx *, int, hipStream_t);
template hipError_t launch_thermal<float>(wfx_ctx *, hipStream_t);
template hipError_t launch_thermal_expansion<float>(wfx_ctx *, hipStream_t);
template hipError_t launch_fused_steps<float>(wfx_ctx *, int, double *, hipStream_t, hipEvent_t *);
template hipError_t launch_fused_pack<float>(wfx_ctx *, hipStream_t);
template hipError_t launch_fused_unpack<float>(wfx_ctx *, hipStream_t);
template hipError_t launch_cfl_begin<float>(wfx_ctx *, hipStream_t, bool);
}  // namespace wfx
