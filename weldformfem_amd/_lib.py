"""ctypes binding of the wfx C-ABI (include/wfx.h).

The library is loaded from this package directory (built in-tree by
weldformfem_amd/build.py).  There is no fallback: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwfx.so")

WFX_FP64 = 0
WFX_FP32_MIXED = 1
WFX_BILINEAR = 0
WFX_HOLLOMON = 1
WFX_JOHNSON_COOK = 2
WFX_GMT = 3

_vp, _i, _d, _cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); every function returns an int status
SIGNATURES = {
    "wfx_last_error": (_cp, []),
    "wfx_version": (_i, []),
    "wfx_create": (_i, [ctypes.POINTER(_vp), _i]),
    "wfx_destroy": (_i, [_vp]),
    "wfx_add_box_length": (_i, [_vp, _pd, _pd, _d, _i, _i]),
    "wfx_set_mesh": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "wfx_get_sizes": (_i, [_vp, _pi, _pi, _pi]),
    "wfx_set_density": (_i, [_vp, _d]),
    "wfx_set_material": (_i, [_vp, _i, _d, _d, _d, _d, _d, _d]),
    "wfx_set_elem_sy0": (_i, [_vp, _vp, _i]),
    "wfx_set_material_jc": (_i, [_vp, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d]),
    "wfx_set_material_gmt": (_i, [_vp, _d, _d, _d, _d, _vp, _vp]),
    "wfx_set_temperature": (_i, [_vp, _d]),
    "wfx_set_elem_temperature": (_i, [_vp, _vp, _i]),
    "wfx_set_max_strain_rate": (_i, [_vp, _d]),
    "wfx_set_cfl": (_i, [_vp, _d]),
    "wfx_set_thermal": (_i, [_vp, _i, _d, _d, _d]),
    "wfx_set_rigid_mesh": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "wfx_set_contact_params": (_i, [_vp, _d, _d, _d]),
    "wfx_set_contact_heat": (_i, [_vp, _d, _d]),
    "wfx_set_rigid_velocity": (_i, [_vp, _vp]),
    "wfx_contact_force": (_i, [_vp, _i, _vp]),
    "wfx_calc_ext_face_areas": (_i, [_vp, _vp]),
    "wfx_calc_contact_forces": (_i, [_vp, _vp]),
    "wfx_rigid_update": (_i, [_vp, _vp]),
    "wfx_thermal_calcs": (_i, [_vp, _vp]),
    "wfx_set_thermal_expansion": (_i, [_vp, _d]),
    "wfx_calc_thermal_expansion": (_i, [_vp, _vp]),
    "wfx_get_dt": (_i, [_vp, _vp]),
    "wfx_calc_cfl_dt": (_i, [_vp, _vp, _vp]),
    "wfx_dt_history": (_i, [_vp, _vp, _i]),
    "wfx_add_bc_vel_node": (_i, [_vp, _i, _i, _d]),
    "wfx_allocate_bcs": (_i, [_vp]),
    "wfx_set_dt": (_i, [_vp, _d]),
    "wfx_set_ch": (_i, [_vp, _d, _d, _d]),
    "wfx_set_param": (_i, [_vp, _cp, _d]),
    "wfx_get_param": (_i, [_vp, _cp, _pd]),
    "wfx_set_reaction_nodes": (_i, [_vp, _vp, _i]),
    "wfx_calc_min_edge_length": (_i, [_vp, _pd, _vp]),
    "wfx_init": (_i, [_vp, _vp]),
    "wfx_update_prediction": (_i, [_vp, _vp]),
    "wfx_impose_bcv": (_i, [_vp, _i, _vp]),
    "wfx_calc_elem_jacobians": (_i, [_vp, _vp]),
    "wfx_calc_elem_vol": (_i, [_vp, _vp]),
    "wfx_calc_elem_density": (_i, [_vp, _vp]),
    "wfx_calc_nodal_vol": (_i, [_vp, _vp]),
    "wfx_calc_nodal_mass_from_vol": (_i, [_vp, _vp]),
    "wfx_calc_elem_strain_rates": (_i, [_vp, _vp]),
    "wfx_calc_elem_pressure": (_i, [_vp, _vp]),
    "wfx_calc_stress_strain": (_i, [_vp, _d, _vp]),
    "wfx_calc_artificial_viscosity": (_i, [_vp, _vp]),
    "wfx_calc_elem_forces": (_i, [_vp, _vp]),
    "wfx_calc_elem_hourglass_forces": (_i, [_vp, _vp]),
    "wfx_assembly_forces": (_i, [_vp, _vp]),
    "wfx_calc_accel": (_i, [_vp, _vp]),
    "wfx_impose_bca": (_i, [_vp, _i, _vp]),
    "wfx_update_correction_accvel": (_i, [_vp, _vp]),
    "wfx_update_correction_pos": (_i, [_vp, _vp]),
    "wfx_step_api": (_i, [_vp, _vp]),
    "wfx_run": (_i, [_vp, _i, _vp, _vp]),
    "wfx_step_fused": (_i, [_vp, _i, _vp, _vp]),
    "wfx_fused_phase": (_i, [_vp, _i, _i, _vp]),
    "wfx_set_reaction_groups": (_i, [_vp, _vp, _i, _vp, _i]),
    "wfx_reaction_group_count": (_i, [_vp, _pi]),
    "wfx_set_elem_tangent": (_i, [_vp, _vp, _i]),
    "wfx_autotune": (_i, [_vp, _i, _vp, _vp]),
    "wfx_set_block_sizes": (_i, [_vp, _vp]),
    "wfx_graph_capture": (_i, [_vp, _i, _i, _vp]),
    "wfx_graph_launch": (_i, [_vp, _vp, _vp]),
    "wfx_adjoint_forward": (_i, [_vp, _i, _i, _vp, _vp]),
    "wfx_adjoint_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wfx_adjoint_release": (_i, [_vp]),
    "wfx_parse_lsdyna_k": (_i, [_cp, _pi, _pi, _pi, _vp, _vp]),
    "wfx_read_lsdyna_k": (_i, [_vp, _cp]),
    "wfx_write_vtk": (_i, [_vp, _cp, _vp]),
    "wfx_write_vtk_async": (_i, [_vp, _cp, _vp]),
    "wfx_vtk_wait": (_i, [_vp]),
    "wfx_reaction_force": (_i, [_vp, _i, _pd, _vp]),
    "wfx_energies": (_i, [_vp, _pd, _pd, _vp]),
    "wfx_time_fused_phases": (_i, [_vp, _i, _pd, _vp]),
    "wfx_energies_noalloc": (_i, [_vp, _pd, _pd, _vp]),
    "wfx_nonfinite_count": (_i, [_vp, ctypes.POINTER(ctypes.c_long)]),
    "wfx_inverted_count": (_i, [_vp, ctypes.POINTER(ctypes.c_long)]),
    "wfx_stream_copy": (_i, [_vp, _vp, ctypes.c_long, _vp]),
    "wfx_field_count": (_i, [_vp, _cp, ctypes.POINTER(ctypes.c_long), _pi]),
    "wfx_get_field": (_i, [_vp, _cp, _vp, ctypes.c_long]),
    "wfx_set_field": (_i, [_vp, _cp, _vp, ctypes.c_long]),
}

_lib = None


class WfxError(RuntimeError):
    pass


def load():
    """Load libwfx.so (raises if it has not been built).

    PyTorch-ROCm ships its own copy of the HIP runtime.  torch is imported
    first so that libwfx binds to that same copy: with two runtimes in one
    process, stream handles made by torch are not valid in the library and
    whichever runtime starts second may find no device."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WfxError(f"wfx library not built: {LIB_PATH} (run weldformfem_amd/build.py)")
        try:
            import torch  # noqa: F401  (one HIP runtime per process)
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != 0:
        msg = load().wfx_last_error()
        raise WfxError(f"wfx error {rc}: {msg.decode() if msg else ''}")
    return rc
