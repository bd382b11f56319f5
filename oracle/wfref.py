"""ctypes front end of the oracle (oracle/libwfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the results to compare against and by bench.py's cpu_baseline leg.  The
product path never uses it.  Arrays come back in the reference's own layout (Domain_d: AoS inside a
field, element tensors as [xx,yy,zz,xy,yz,xz], element forces as [e][a][d]).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libwfref.so")
_lib = None

BILINEAR = 0
HOLLOMON = 1
JOHNSON_COOK = 2
GMT = 3


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, d, dp, ip = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p
        sig = {
            "wfref_create": (vp, []),
            "wfref_destroy": (None, [vp]),
            "wfref_add_box_length": (i, [vp, dp, dp, d, i, i]),
            "wfref_create_from_mesh": (i, [vp, i, i, i, dp, ip]),
            "wfref_set_density": (None, [vp, d]),
            "wfref_set_material": (None, [vp, i, d, d, d, d, d, d]),
            "wfref_set_elem_sy0": (None, [vp, dp]),
            "wfref_set_material_jc": (None, [vp, d, d, d, d, d, d, d, d, d, d, d]),
            "wfref_set_thermal": (None, [vp, i, d, d, d]),
            "wfref_set_rigid_mesh": (i, [vp, i, dp, dp, i, ip, ip]),
            "wfref_calc_ext_face_areas": (None, [vp]),
            "wfref_calc_elem_length": (None, [vp]),
            "wfref_calc_contact_forces": (None, [vp]),
            "wfref_rigid_update": (None, [vp]),
            "wfref_contact_force": (d, [vp, i]),
            "wfref_thermal_calcs": (None, [vp]),
            "wfref_set_thermal_expansion": (None, [vp, d]),
            "wfref_calc_thermal_expansion": (None, [vp]),
            "wfref_set_material_gmt": (None, [vp, d, d, d, d, dp, dp]),
            "wfref_eff_strain_rate": (d, [dp]),
            "wfref_yield_stress": (d, [vp, i, d, d]),
            "wfref_tangent_modulus": (d, [vp, i, d, d]),
            "wfref_add_bc_vel_node": (None, [vp, i, i, d]),
            "wfref_set_reaction_nodes": (None, [vp, i, ip]),
            "wfref_set_ch": (None, [vp, d, d, d]),
            "wfref_calc_min_edge_length": (d, [vp]),
            "wfref_init": (None, [vp]),
            "wfref_step": (None, [vp, dp]),
            "wfref_run": (None, [vp, i, dp]),
            "wfref_energies": (None, [vp, dp, dp]),
            "wfref_reaction_force": (d, [vp, i]),
            "wfref_field": (vp, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long),
                                 ctypes.POINTER(ctypes.c_int)]),
            "wfref_get_int": (i, [vp, ctypes.c_char_p]),
            "wfref_get_double": (d, [vp, ctypes.c_char_p]),
            "wfref_set_double": (i, [vp, ctypes.c_char_p, d]),
            "wfref_calc_stress_strain": (None, [vp, d]),
            "wfref_impose_bcv": (None, [vp, i]),
            "wfref_impose_bca": (None, [vp, i]),
            "wfref_num_threads": (i, []),
        }
        for name in ["update_prediction", "calc_elem_jacobians", "calc_elem_vol",
                     "calc_elem_density", "calc_nodal_vol", "calc_nodal_mass_from_vol",
                     "calc_elem_strain_rates", "calc_elem_pressure",
                     "calc_artificial_viscosity", "calc_elem_forces",
                     "calc_elem_hourglass_forces", "assembly_forces",
                     "scrub_nonfinite_fi", "calc_accel", "update_correction_accvel",
                     "update_correction_pos"]:
            sig["wfref_" + name] = (None, [vp])
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def num_threads():
    return lib().wfref_num_threads()


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class RefDomain:
    """Oracle domain; method names follow MetFEM::Domain_d."""

    def __init__(self):
        self._L = lib()
        self._d = self._L.wfref_create()

    def __del__(self):
        try:
            if self._d:
                self._L.wfref_destroy(self._d)
                self._d = None
        except Exception:
            pass

    # -- mesh -------------------------------------------------------------
    def AddBoxLength(self, V, L, r, red_int=True, tritetra=False):
        Va = np.ascontiguousarray(V, dtype=np.float64)
        La = np.ascontiguousarray(L, dtype=np.float64)
        rc = self._L.wfref_add_box_length(self._d, _ptr(Va), _ptr(La), float(r), int(red_int), int(tritetra))
        if rc:
            raise ValueError("AddBoxLength: unsupported arguments")

    def CreateFromMesh(self, x, elnod, nodxelem):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        el = np.ascontiguousarray(elnod, dtype=np.int32).reshape(-1)
        N = x.size // 3
        E = el.size // nodxelem
        rc = self._L.wfref_create_from_mesh(self._d, N, E, nodxelem, _ptr(x), _ptr(el))
        if rc:
            raise ValueError("CreateFromMesh: unsupported element")

    # -- setup ------------------------------------------------------------
    def setDensity(self, rho):
        self._L.wfref_set_density(self._d, float(rho))

    def setMaterial(self, model, E, nu, rho0, sy0, p1=0.0, p2=0.0):
        self._L.wfref_set_material(self._d, int(model), E, nu, rho0, sy0, p1, p2)

    def setMaterialJC(self, E, nu, rho0, A, B, n, C, eps_0, m, T_m, T_t):
        self._L.wfref_set_material_jc(self._d, E, nu, rho0, A, B, n, C, eps_0, m, T_m, T_t)

    def setMaterialGMT(self, E, nu, rho0, sy0, c, ranges):
        ca = np.ascontiguousarray(c, dtype=np.float64)
        ra = np.ascontiguousarray(ranges, dtype=np.float64)
        assert ca.size == 8 and ra.size == 6
        self._L.wfref_set_material_gmt(self._d, E, nu, rho0, sy0, _ptr(ca), _ptr(ra))

    def setTemperature(self, T):
        """uniform temperature (elements and nodes, as setTemp), or one
        value per element"""
        a = np.asarray(T, dtype=np.float64)
        t = np.broadcast_to(a, (self.elem_count,))
        self.setField("temp", np.ascontiguousarray(t))
        if a.ndim == 0:
            self.setField("T", np.full(self.node_count, float(a)))

    def setCFL(self, cfl_factor):
        """per-step CFL time step (0: fixed dt)"""
        self.set("cfl", cfl_factor)

    def getDT(self):
        return self._L.wfref_get_double(self._d, b"dt")

    def setThermal(self, on=True, k_T=50.0, cp_T=460.0, plheat_frac=0.9):
        """thermal coupling (nodal temperatures in field "T")"""
        self._L.wfref_set_thermal(self._d, int(on), float(k_T), float(cp_T), float(plheat_frac))

    def ThermalCalcs(self):
        self._L.wfref_thermal_calcs(self._d)

    def setThermalExpansion(self, exp_T):
        """material thermal expansion coefficient (thermalExp)"""
        self._L.wfref_set_thermal_expansion(self._d, float(exp_T))

    def calcThermalExpansion(self):
        self._L.wfref_calc_thermal_expansion(self._d)

    def setRigidMesh(self, nodes, tris, vel=None, mesh_id=None):
        """rigid triangle surface for penalty contact (tetrahedral body)"""
        x = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1)
        v = np.zeros_like(x) if vel is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(vel, dtype=np.float64), (x.size // 3, 3))).reshape(-1)
        m = np.zeros(t.size // 3, np.int32) if mesh_id is None else np.ascontiguousarray(mesh_id, dtype=np.int32)
        rc = self._L.wfref_set_rigid_mesh(self._d, x.size // 3, _ptr(x), _ptr(v), t.size // 3, _ptr(t), _ptr(m))
        if rc:
            raise ValueError("rigid mesh: needs a tetrahedral body and a non-empty surface")

    def setContactParams(self, penalty_factor=0.1, mu_static=0.0, mu_dynamic=0.0):
        self.set("contPF", penalty_factor)
        self.set("mu_sta", mu_static)
        self.set("mu_dyn", mu_dynamic)

    def setContactHeat(self, heat_cond, die_temp):
        """contact heat conduction: heatCondCoeff and dieTemp (main.C:718-719)"""
        self.set("heat_cond", heat_cond)
        self.set("T_const", die_temp)

    def CalcExtFaceAreas(self):
        self._L.wfref_calc_ext_face_areas(self._d)

    def calcElemLength(self):
        self._L.wfref_calc_elem_length(self._d)

    def CalcContactForces(self):
        self._L.wfref_calc_contact_forces(self._d)

    def rigidUpdate(self):
        self._L.wfref_rigid_update(self._d)

    def contactForce(self, mesh_id=0):
        return self._L.wfref_contact_force(self._d, int(mesh_id))

    def setMaxStrainRate(self, edot):
        self.set("max_edot", edot)

    def yieldStress(self, e, ep, edot):
        return self._L.wfref_yield_stress(self._d, int(e), float(ep), float(edot))

    def tangentModulus(self, e, ep, edot):
        return self._L.wfref_tangent_modulus(self._d, int(e), float(ep), float(edot))

    @staticmethod
    def effStrainRate(D6):
        a = np.ascontiguousarray(D6, dtype=np.float64)
        return lib().wfref_eff_strain_rate(_ptr(a))

    def setElemSy0(self, sy0):
        a = np.ascontiguousarray(sy0, dtype=np.float64)
        assert a.size == self.elem_count
        self._L.wfref_set_elem_sy0(self._d, _ptr(a))

    def AddBCVelNode(self, node, dim, val):
        self._L.wfref_add_bc_vel_node(self._d, int(node), int(dim), float(val))

    def setReactionNodes(self, nodes):
        a = np.ascontiguousarray(nodes, dtype=np.int32)
        self._L.wfref_set_reaction_nodes(self._d, a.size, _ptr(a))

    def setCH(self, alpha, beta, gamma):
        self._L.wfref_set_ch(self._d, alpha, beta, gamma)

    def set(self, name, value):
        if self._L.wfref_set_double(self._d, name.encode(), float(value)):
            raise KeyError(name)

    def SetDT(self, dt):
        self.set("dt", dt)

    def get(self, name):
        return self._L.wfref_get_double(self._d, name.encode())

    def calcMinEdgeLength(self):
        return self._L.wfref_calc_min_edge_length(self._d)

    # -- sizes / fields ---------------------------------------------------
    @property
    def node_count(self):
        return self._L.wfref_get_int(self._d, b"node_count")

    @property
    def elem_count(self):
        return self._L.wfref_get_int(self._d, b"elem_count")

    @property
    def nodxelem(self):
        return self._L.wfref_get_int(self._d, b"nodxelem")

    @property
    def nonfinite_fi(self):
        return self._L.wfref_get_int(self._d, b"nonfinite_fi")

    @property
    def inverted(self):
        return self._L.wfref_get_int(self._d, b"inverted")

    def field(self, name, copy=True):
        cnt = ctypes.c_long()
        isint = ctypes.c_int()
        p = self._L.wfref_field(self._d, name.encode(), ctypes.byref(cnt), ctypes.byref(isint))
        if not p:
            raise KeyError(name)
        ct = ctypes.c_int if isint.value else ctypes.c_double
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(cnt.value,))
        return arr.copy() if copy else arr

    def setField(self, name, values):
        dst = self.field(name, copy=False)
        dst[:] = np.asarray(values, dtype=dst.dtype).reshape(-1)

    # -- solver -----------------------------------------------------------
    def init(self):
        self._L.wfref_init(self._d)

    def step(self):
        r = ctypes.c_double()
        self._L.wfref_step(self._d, ctypes.byref(r))
        return r.value

    def run(self, nsteps):
        hist = np.zeros(max(nsteps, 1), dtype=np.float64)
        self._L.wfref_run(self._d, int(nsteps), _ptr(hist))
        return hist[:nsteps]

    def energies(self):
        ek, de = ctypes.c_double(), ctypes.c_double()
        self._L.wfref_energies(self._d, ctypes.byref(ek), ctypes.byref(de))
        return ek.value, de.value

    def __getattr__(self, name):
        # per-step API by reference name, e.g. calcElemJAndDerivatives()
        table = {
            "UpdatePrediction": "update_prediction",
            "calcElemJAndDerivatives": "calc_elem_jacobians",
            "CalcElemVol": "calc_elem_vol",
            "calcElemDensity": "calc_elem_density",
            "CalcNodalVol": "calc_nodal_vol",
            "CalcNodalMassFromVol": "calc_nodal_mass_from_vol",
            "calcElemStrainRates": "calc_elem_strain_rates",
            "calcElemPressure": "calc_elem_pressure",
            "calcArtificialViscosity": "calc_artificial_viscosity",
            "calcElemForces": "calc_elem_forces",
            "calcElemHourglassForces": "calc_elem_hourglass_forces",
            "assemblyForces": "assembly_forces",
            "scrubNonFiniteForces": "scrub_nonfinite_fi",
            "calcAccel": "calc_accel",
            "UpdateCorrectionAccVel": "update_correction_accvel",
            "UpdateCorrectionPos": "update_correction_pos",
        }
        if name in table:
            fn = getattr(self._L, "wfref_" + table[name])
            return lambda: fn(self._d)
        if name == "CalcStressStrain":
            return lambda dt: self._L.wfref_calc_stress_strain(self._d, float(dt))
        if name == "ImposeBCV":
            return lambda dim: self._L.wfref_impose_bcv(self._d, int(dim))
        if name == "ImposeBCA":
            return lambda dim: self._L.wfref_impose_bca(self._d, int(dim))
        raise AttributeError(name)
