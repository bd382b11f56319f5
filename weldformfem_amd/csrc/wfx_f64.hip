// FP64 instantiation of the fused step (F1..F4 and their helpers), built
// with floating-point contraction: the fused path's sums already differ from
// the reference's order (per-cluster nodal sums, per-element hourglass), so
// multiply-adds take one rounding instead of two; the per-step kernels
// (wfx_k64.hip) keep the reference's roundings.
#include "wfx_kernels.hpp"
#include "wfx_fused.hpp"
#include "wfx_launch.hpp"

namespace wfx {
template hipError_t launch_vsum0<double>(wfx_ctx *, hipStream_t);
template hipError_t launch_fused_phase<double>(wfx_ctx *, int, int, hipStream_t, int);
template hipError_t launch_fused_steps<double>(wfx_ctx *, int, double *, hipStream_t, hipEvent_t *);
template hipError_t launch_fused_pack<double>(wfx_ctx *, hipStream_t);
template hipError_t launch_fused_unpack<double>(wfx_ctx *, hipStream_t);
template hipError_t launch_cfl_begin<double>(wfx_ctx *, hipStream_t, bool);
}  // namespace wfx
