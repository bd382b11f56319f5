"""weldformfem_amd: MI355X-native explicit updated-Lagrangian time step of
WeldFormFEM's Domain_d, as HIP kernels for gfx950 behind a C-ABI
(include/wfx.h, libwfx.so in this directory)."""
from .domain import (Domain, WFX_BILINEAR, WFX_FP32_MIXED, WFX_FP64, WFX_GMT, WFX_HOLLOMON,  # noqa: F401
                     WFX_JOHNSON_COOK)
