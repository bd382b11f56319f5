// FP32-mixed instantiation of the per-step element kernels.
#include "wfx_kernels.hpp"
#include "wfx_launch.hpp"

namespace wfx {
WFX_DEFINE_LAUNCHERS(float)
}  // namespace wfx
