"""Python mirror of MetFEM::Domain_d over the wfx C-ABI.

Method names, argument meaning and call order follow the reference
(include/common/Domain_d.h:238-830, src/explicit/Solver_explicit.C) so a
reference solve loop reads the same here.  Every call goes to the HIP
library; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import WFX_BILINEAR, WFX_FP32_MIXED, WFX_FP64, WFX_GMT, WFX_HOLLOMON, WFX_JOHNSON_COOK, check

__all__ = ["Domain", "WFX_FP64", "WFX_FP32_MIXED", "WFX_BILINEAR", "WFX_HOLLOMON", "WFX_JOHNSON_COOK", "WFX_GMT"]


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Domain:
    """One explicit domain on the GPU.  ``stream`` is a hipStream_t handle
    (int, e.g. torch.cuda.current_stream().cuda_stream); 0 = default stream."""

    def __init__(self, precision=WFX_FP64, stream=0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        check(self._L.wfx_create(ctypes.byref(h), int(precision)))
        self._h = h
        self.precision = precision
        self.stream = ctypes.c_void_p(stream or None)
        self.law = WFX_BILINEAR  # the hardening law last set

    def close(self):
        if getattr(self, "_h", None):
            self._L.wfx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- mesh --------------------------------------------------------------
    def AddBoxLength(self, V, L, r, red_int=True, tritetra=False):
        Va = np.ascontiguousarray(V, dtype=np.float64)
        La = np.ascontiguousarray(L, dtype=np.float64)
        check(self._L.wfx_add_box_length(self._h, _dptr(Va), _dptr(La), float(r), int(red_int), int(tritetra)))

    def CreateFromMesh(self, x, elnod, nodxelem):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        el = np.ascontiguousarray(elnod, dtype=np.int32).reshape(-1)
        check(self._L.wfx_set_mesh(self._h, x.size // 3, el.size // nodxelem, int(nodxelem),
                                   x.ctypes.data, el.ctypes.data))

    def readLSDyna(self, path):
        """mesh from an LS-DYNA keyword file (*NODE, *ELEMENT_SOLID)"""
        check(self._L.wfx_read_lsdyna_k(self._h, str(path).encode()))

    def writeVTK(self, path):
        """legacy VTK frame of the current state"""
        check(self._L.wfx_write_vtk(self._h, str(path).encode(), self.stream))

    def writeVTKAsync(self, path):
        """the same frame written by a host thread while stepping continues"""
        check(self._L.wfx_write_vtk_async(self._h, str(path).encode(), self.stream))

    def waitVTK(self):
        """waits for the pending asynchronous frames"""
        check(self._L.wfx_vtk_wait(self._h))

    def sizes(self):
        N, E, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self._L.wfx_get_sizes(self._h, ctypes.byref(N), ctypes.byref(E), ctypes.byref(n)))
        return N.value, E.value, n.value

    @property
    def node_count(self):
        return self.sizes()[0]

    @property
    def elem_count(self):
        return self.sizes()[1]

    @property
    def nodxelem(self):
        return self.sizes()[2]

    # ---- set-up ------------------------------------------------------------
    def setDensity(self, rho):
        check(self._L.wfx_set_density(self._h, float(rho)))

    def setMaterial(self, model, E, nu, rho0, sy0, p1=0.0, p2=0.0):
        """bilinear (p1 = Et) or Hollomon (p1 = K, p2 = m)"""
        check(self._L.wfx_set_material(self._h, int(model), E, nu, rho0, sy0, p1, p2))
        self.law = int(model)

    def setMaterialJC(self, E, nu, rho0, A, B, n, C, eps_0, m, T_m, T_t):
        """Johnson-Cook law (Material.cuh:191-203, :378-411); A is the
        per-element initial yield"""
        check(self._L.wfx_set_material_jc(self._h, E, nu, rho0, A, B, n, C, eps_0, m, T_m, T_t))
        self.law = WFX_JOHNSON_COOK

    def setMaterialGMT(self, E, nu, rho0, sy0, c, ranges):
        """GMT law (Material.cuh:271-291, :412-473); c = (n1, n2, C1, C2, m1,
        m2, I1, I2), ranges = (e_min, e_max, er_min, er_max, T_min, T_max)"""
        ca = np.ascontiguousarray(c, dtype=np.float64)
        ra = np.ascontiguousarray(ranges, dtype=np.float64)
        if ca.size != 8 or ra.size != 6:
            raise ValueError("GMT needs 8 constants and 6 range bounds")
        check(self._L.wfx_set_material_gmt(self._h, E, nu, rho0, sy0, ca.ctypes.data, ra.ctypes.data))
        self.law = WFX_GMT

    def setTemperature(self, T):
        """uniform temperature (Domain_d::setTemp), or one value per element"""
        a = np.asarray(T, dtype=np.float64)
        if a.ndim == 0:
            check(self._L.wfx_set_elem_temperature(self._h, None, 0))
            check(self._L.wfx_set_temperature(self._h, float(a)))
        else:
            a = np.ascontiguousarray(a.reshape(-1))
            check(self._L.wfx_set_elem_temperature(self._h, a.ctypes.data, a.size))

    def setRigidMesh(self, nodes, tris, vel=None, mesh_id=None):
        """rigid triangle surfaces for penalty contact (TriMesh_d + setContactOn);
        the body must be tetrahedral"""
        x = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1)
        v = None
        if vel is not None:
            v = np.ascontiguousarray(np.broadcast_to(np.asarray(vel, dtype=np.float64), (x.size // 3, 3))).reshape(-1)
        m = None if mesh_id is None else np.ascontiguousarray(mesh_id, dtype=np.int32)
        check(self._L.wfx_set_rigid_mesh(self._h, x.size // 3, x.ctypes.data, None if v is None else v.ctypes.data,
                                         t.size // 3, t.ctypes.data, None if m is None else m.ctypes.data))

    def setContactParams(self, penalty_factor=0.1, mu_static=0.0, mu_dynamic=0.0):
        check(self._L.wfx_set_contact_params(self._h, float(penalty_factor), float(mu_static), float(mu_dynamic)))

    def setContactHeat(self, heat_cond, die_temp):
        """contact heat conduction: heatCondCoeff and dieTemp (main.C:718-719)"""
        check(self._L.wfx_set_contact_heat(self._h, float(heat_cond), float(die_temp)))

    def setRigidVelocity(self, vel, node_count):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(vel, dtype=np.float64), (int(node_count), 3))).reshape(-1)
        check(self._L.wfx_set_rigid_velocity(self._h, v.ctypes.data))

    def contactForce(self, mesh_id=0):
        v = ctypes.c_double()
        check(self._L.wfx_contact_force(self._h, int(mesh_id), ctypes.byref(v)))
        return v.value

    def CalcExtFaceAreas(self): self._call(self._L.wfx_calc_ext_face_areas)
    def CalcContactForces(self): self._call(self._L.wfx_calc_contact_forces)
    def rigidUpdate(self): self._call(self._L.wfx_rigid_update)

    def setThermal(self, on=True, k_T=50.0, cp_T=460.0, plheat_frac=0.9):
        """thermal coupling (setThermalOn, thermalCond, thermalHeatCap,
        plHeatFrac); nodal temperatures in the field "T"
        """
        check(self._L.wfx_set_thermal(self._h, int(on), float(k_T), float(cp_T), float(plheat_frac)))

    def ThermalCalcs(self):
        self._call(self._L.wfx_thermal_calcs)

    def setThermalExpansion(self, exp_T):
        """material thermalExp (Material_::exp_T): with thermal coupling,
        every step runs calcThermalExpansion after the strain rates"""
        check(self._L.wfx_set_thermal_expansion(self._h, float(exp_T)))

    def calcThermalExpansion(self):
        self._call(self._L.wfx_calc_thermal_expansion)

    def setMaxStrainRate(self, edot):
        """Domain_d::m_max_edot: cap of the rate the tangent modulus sees"""
        check(self._L.wfx_set_max_strain_rate(self._h, float(edot)))

    def setElemSy0(self, sy0):
        a = np.ascontiguousarray(sy0, dtype=np.float64)
        check(self._L.wfx_set_elem_sy0(self._h, a.ctypes.data, a.size))

    def AddBCVelNode(self, node, dim, val):
        check(self._L.wfx_add_bc_vel_node(self._h, int(node), int(dim), float(val)))

    def AllocateBCs(self):
        check(self._L.wfx_allocate_bcs(self._h))

    def SetDT(self, dt):
        check(self._L.wfx_set_dt(self._h, float(dt)))

    def setCFL(self, cfl_factor):
        """per-step CFL time step with this factor (Domain_d::setCFL with
        m_fixed_dt false); 0 keeps dt fixed"""
        check(self._L.wfx_set_cfl(self._h, float(cfl_factor)))

    def calcCflDt(self):
        """per-step CFL time step of the current state; the following
        per-step calls use it"""
        v = ctypes.c_double()
        check(self._L.wfx_calc_cfl_dt(self._h, ctypes.byref(v), self.stream))
        return v.value

    def getDT(self):
        v = ctypes.c_double()
        check(self._L.wfx_get_dt(self._h, ctypes.byref(v)))
        return v.value

    def dtHistory(self, count):
        """dt of each step of the last run with the per-step CFL time step"""
        out = np.zeros(int(count), dtype=np.float64)
        check(self._L.wfx_dt_history(self._h, out.ctypes.data, int(count)))
        return out

    def setCH(self, alpha, beta, gamma):
        check(self._L.wfx_set_ch(self._h, alpha, beta, gamma))

    def set(self, name, value):
        check(self._L.wfx_set_param(self._h, name.encode(), float(value)))

    def get(self, name):
        v = ctypes.c_double()
        check(self._L.wfx_get_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def setReactionNodes(self, nodes):
        a = np.ascontiguousarray(nodes, dtype=np.int32)
        check(self._L.wfx_set_reaction_nodes(self._h, a.ctypes.data, a.size))

    def calcMinEdgeLength(self):
        v = ctypes.c_double()
        check(self._L.wfx_calc_min_edge_length(self._h, ctypes.byref(v), self.stream))
        return v.value

    # ---- solver ------------------------------------------------------------
    def init(self):
        check(self._L.wfx_init(self._h, self.stream))

    def _call(self, fn, *args):
        check(fn(self._h, *args, self.stream))

    def UpdatePrediction(self): self._call(self._L.wfx_update_prediction)
    def ImposeBCV(self, dim): self._call(self._L.wfx_impose_bcv, int(dim))
    def ImposeBCA(self, dim): self._call(self._L.wfx_impose_bca, int(dim))
    def calcElemJAndDerivatives(self): self._call(self._L.wfx_calc_elem_jacobians)
    def CalcElemVol(self): self._call(self._L.wfx_calc_elem_vol)
    def calcElemDensity(self): self._call(self._L.wfx_calc_elem_density)
    def CalcNodalVol(self): self._call(self._L.wfx_calc_nodal_vol)
    def CalcNodalMassFromVol(self): self._call(self._L.wfx_calc_nodal_mass_from_vol)
    def calcElemStrainRates(self): self._call(self._L.wfx_calc_elem_strain_rates)
    def calcElemPressure(self): self._call(self._L.wfx_calc_elem_pressure)
    def CalcStressStrain(self, dt): self._call(self._L.wfx_calc_stress_strain, float(dt))
    def calcArtificialViscosity(self): self._call(self._L.wfx_calc_artificial_viscosity)
    def calcElemForces(self): self._call(self._L.wfx_calc_elem_forces)
    def calcElemHourglassForces(self): self._call(self._L.wfx_calc_elem_hourglass_forces)
    def assemblyForces(self): self._call(self._L.wfx_assembly_forces)
    def calcAccel(self): self._call(self._L.wfx_calc_accel)
    def UpdateCorrectionAccVel(self): self._call(self._L.wfx_update_correction_accvel)
    def UpdateCorrectionPos(self): self._call(self._L.wfx_update_correction_pos)

    def step(self):
        check(self._L.wfx_step_api(self._h, self.stream))

    def setReactionGroups(self, groups):
        """one reaction node list per group (e.g. per specimen of a batch)"""
        arrs = [np.ascontiguousarray(g, dtype=np.int32).reshape(-1) for g in groups]
        off = np.zeros(len(arrs) + 1, dtype=np.int32)
        off[1:] = np.cumsum([a.size for a in arrs])
        nodes = np.concatenate(arrs) if arrs else np.zeros(0, np.int32)
        check(self._L.wfx_set_reaction_groups(self._h, nodes.ctypes.data, nodes.size, off.ctypes.data,
                                              len(arrs)))

    def reactionGroupCount(self):
        v = ctypes.c_int()
        check(self._L.wfx_reaction_group_count(self._h, ctypes.byref(v)))
        return v.value

    def setElemTangent(self, Et):
        a = np.ascontiguousarray(Et, dtype=np.float64)
        check(self._L.wfx_set_elem_tangent(self._h, a.ctypes.data, a.size))

    def _hist(self, nsteps, raw):
        """reaction history shaped (nsteps,) for one group, (nsteps, groups) else"""
        ng = max(self.reactionGroupCount(), 1)
        if ng == 1:
            return raw[:nsteps]
        return raw[: nsteps * ng].reshape(nsteps, ng)

    def run(self, nsteps, reaction=True):
        ng = max(self.reactionGroupCount(), 1)
        hist = np.zeros(max(int(nsteps), 1) * ng, dtype=np.float64)
        check(self._L.wfx_run(self._h, int(nsteps), hist.ctypes.data if reaction else None, self.stream))
        return self._hist(nsteps, hist)

    def runFused(self, nsteps, reaction=True):
        """nsteps steps through the four-kernel fused path (wfx_step_fused)"""
        ng = max(self.reactionGroupCount(), 1)
        hist = np.zeros(max(int(nsteps), 1) * ng, dtype=np.float64)
        check(self._L.wfx_step_fused(self._h, int(nsteps), hist.ctypes.data if reaction else None,
                                     self.stream))
        return self._hist(nsteps, hist)

    def fusedPhase(self, phase, predict=True):
        """launch one kernel of the fused step (timing and profiling)"""
        check(self._L.wfx_fused_phase(self._h, int(phase), int(predict), self.stream))

    def timeFusedPhases(self, nsteps):
        """nsteps fused steps with HIP events between the four passes: the
        average microseconds of each pass, in the order they run"""
        us = (ctypes.c_double * 4)()
        check(self._L.wfx_time_fused_phases(self._h, int(nsteps), us, self.stream))
        return list(us)

    def autotune(self, reps=5):
        """pick the block size of each fused kernel by timing; returns the four sizes"""
        out = (ctypes.c_int * 4)()
        check(self._L.wfx_autotune(self._h, int(reps), out, self.stream))
        return list(out)

    def setBlockSizes(self, sizes):
        a = (ctypes.c_int * 4)(*[int(s) for s in sizes])
        check(self._L.wfx_set_block_sizes(self._h, a))

    def graphCapture(self, nsteps, reaction=True):
        """capture nsteps fused steps into a HIP graph on this domain's stream"""
        check(self._L.wfx_graph_capture(self._h, int(nsteps), int(bool(reaction)), self.stream))
        self._graph_steps = int(nsteps)

    def graphLaunch(self, reaction=True):
        if not reaction:
            check(self._L.wfx_graph_launch(self._h, None, self.stream))
            return None
        ng = max(self.reactionGroupCount(), 1)
        hist = np.zeros(self._graph_steps * ng, dtype=np.float64)
        check(self._L.wfx_graph_launch(self._h, hist.ctypes.data, self.stream))
        return self._hist(self._graph_steps, hist)

    # ---- adjoint ---------------------------------------------------------
    def adjointForward(self, nsteps, checkpoint_every=50):
        """forward run recorded for adjointBackward; returns the reaction history"""
        ng = max(self.reactionGroupCount(), 1)
        hist = np.zeros(int(nsteps) * ng, dtype=np.float64)
        check(self._L.wfx_adjoint_forward(self._h, int(nsteps), int(checkpoint_every), hist.ctypes.data,
                                          self.stream))
        self._adj_steps = int(nsteps)
        return self._hist(nsteps, hist)

    def adjointBackward(self, hist_bar, per_element=False):
        """dL/d(sigma_y0, Et, hg_coeff) (bilinear law), dL/d(K, m,
        hg_coeff) (Hollomon), dL/d(A, B, hg_coeff) (Johnson-Cook) or dL/d(C1,
        n2, hg_coeff) (GMT) for the cotangent hist_bar = dL/dR of the recorded
        history (shape as adjointForward returned)"""
        hb = np.ascontiguousarray(hist_bar, dtype=np.float64).reshape(-1)
        ng = max(self.reactionGroupCount(), 1)
        if hb.size != self._adj_steps * ng:
            raise ValueError(f"hist_bar needs {self._adj_steps * ng} values, got {hb.size}")
        g = np.zeros(3)
        E = self.elem_count
        el = [np.zeros(E) for _ in range(3)] if per_element else [None, None, None]
        check(self._L.wfx_adjoint_backward(self._h, hb.ctypes.data, g.ctypes.data,
                                           *[(a.ctypes.data if a is not None else None) for a in el],
                                           self.stream))
        names = {WFX_HOLLOMON: ("K", "m", "hg_coeff"),
                 WFX_JOHNSON_COOK: ("A", "B", "hg_coeff"),
                 WFX_GMT: ("C1", "n2", "hg_coeff")}.get(self.law, ("sy0", "Et", "hg_coeff"))
        out = dict(zip(names, g))
        if per_element:
            out.update({f"{k}_elem": a for k, a in zip(names, el)})
        return out

    def adjointRelease(self):
        check(self._L.wfx_adjoint_release(self._h))

    def reactionForce(self, dim=2):
        ng = max(self.reactionGroupCount(), 1)
        v = np.zeros(ng, dtype=np.float64)
        check(self._L.wfx_reaction_force(self._h, int(dim), _dptr(v), self.stream))
        return float(v[0]) if ng == 1 else v

    def energies(self):
        ek, de = ctypes.c_double(), ctypes.c_double()
        check(self._L.wfx_energies(self._h, ctypes.byref(ek), ctypes.byref(de), self.stream))
        return ek.value, de.value

    def nonfiniteCount(self):
        v = ctypes.c_long()
        check(self._L.wfx_nonfinite_count(self._h, ctypes.byref(v)))
        return v.value

    def invertedCount(self):
        """element Jacobians seen with detJ <= 0 or not finite (one count per
        element and step)"""
        v = ctypes.c_long()
        check(self._L.wfx_inverted_count(self._h, ctypes.byref(v)))
        return v.value

    # ---- fields (reference layout) ------------------------------------------
    def field(self, name):
        cnt, isint = ctypes.c_long(), ctypes.c_int()
        check(self._L.wfx_field_count(self._h, name.encode(), ctypes.byref(cnt), ctypes.byref(isint)))
        out = np.empty(cnt.value, dtype=np.int32 if isint.value else np.float64)
        check(self._L.wfx_get_field(self._h, name.encode(), out.ctypes.data, out.size))
        return out

    def setField(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        check(self._L.wfx_set_field(self._h, name.encode(), a.ctypes.data, a.size))


def parse_lsdyna_k(path):
    """(x[N,3], elnod[E,n], n) of an LS-DYNA keyword mesh, host only"""
    L = _lib.load()
    N, E, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(L.wfx_parse_lsdyna_k(str(path).encode(), ctypes.byref(N), ctypes.byref(E), ctypes.byref(n), None, None))
    x = np.zeros((N.value, 3))
    el = np.zeros((E.value, n.value), dtype=np.int32)
    check(L.wfx_parse_lsdyna_k(str(path).encode(), ctypes.byref(N), ctypes.byref(E), ctypes.byref(n),
                               x.ctypes.data, el.ctypes.data))
    return x, el, n.value
