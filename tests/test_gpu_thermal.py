"""Thermal coupling on the GPU against the oracle.

ThermalCalcs alone on a random state (per-step API, FP64, relative 1e-13),
then 100-step compressions of a Johnson-Cook block whose temperature starts
as a gradient along z: per-step API and fused path, with and without the
per-step CFL time step.  Temperature increments (T - T0), displacements and
plastic strain within 1e-10 relative (FP64) and 1e-3 (FP32-mixed).
A large conductivity makes conduction visible within 100 steps.
Thermal expansion (calcThermalExpansion, Thermal.C:164-178): the kernel on a
random state after ThermalCalcs (1e-13), and the same compressions with a
steel-like coefficient on both paths (1e-10 FP64, 1e-3 FP32-mixed), which
differ from the runs without it.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads

pytestmark = pytest.mark.gpu

JC = (300e6, 500e6, 0.3, 0.014, 1.0, 1.0, 1520.0, 20.0)
K_T, CP = 5.0e4, 460.0


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def jittered_box(nx, ny, nz, h, tet, seed, amp=0.15):
    ref = RefDomain()
    ref.AddBoxLength([0, 0, 0], [nx * h, ny * h, nz * h], h / 2, True, tet)
    x = ref.field("x").reshape(-1, 3)
    x = x + (2.0 * workloads.uniform(seed, x.size).reshape(-1, 3) - 1.0) * amp * h
    return x, ref.field("elnod"), 4 if tet else 8


@pytest.mark.parametrize("tet", [False, True])
def test_thermal_calcs_random_state(tet):
    say(f"case: ThermalCalcs FP64 on a random state, tet={tet}")
    x, el, n = jittered_box(5, 4, 3, 0.02, tet, seed=17)
    g, r = Domain(WFX_FP64), RefDomain()
    for d in (g, r):
        d.CreateFromMesh(x, el, n)
        d.setDensity(7850.0)
        d.setMaterial(0, 200e9, 0.3, 7850.0, 300e6, 1e9, 0.0)
        d.SetDT(1e-4)
        d.init()
        d.setThermal(True, K_T, CP, 0.9)
    rng = np.random.default_rng(5)
    T0 = 20.0 + 800.0 * rng.random(g.node_count)
    q = 1e9 * rng.random(g.elem_count)
    for d in (g, r):
        d.setField("T", T0)
        d.setField("q_plheat", q)
        d.ThermalCalcs()
    e = rel(g.field("T") - T0, r.field("T") - T0)
    assert np.max(np.abs(r.field("T") - T0)) > 1.0
    assert e <= 1e-13, f"T increment: {e:.3e}"


def _hexa16():
    r = RefDomain()
    r.AddBoxLength([0, 0, 0], [0.8, 0.8, 0.8], 0.025)
    bottom, top = workloads.box_layers(16, 16, 16)
    return r.field("x"), r.field("elnod"), 8, bottom, top


def _cyl12():
    x, el, bottom, top = workloads.cylinder_tets(cells=12)
    return x, el, 4, bottom, top


def _pair(mesh, precision=WFX_FP64, cfl=0.0, exp_T=0.0, law="jc"):
    x, el, n, bottom, top = mesh()
    g = Domain(precision)
    g.CreateFromMesh(x, el, n)
    dt = workloads.setup_compression(g, bottom, top, -20.0)
    r = RefDomain()
    r.CreateFromMesh(x, el, n)
    workloads.setup_compression(r, bottom, top, -20.0, dt=dt)
    T0 = 20.0 + 500.0 * np.asarray(x).reshape(-1, 3)[:, 2]
    for d in (g, r):
        if law == "jc":
            d.setMaterialJC(200e9, 0.3, 7850.0, *JC)
        else:  # bilinear, low yield stress: plastic within 100 steps
            d.setMaterial(0, 200e9, 0.3, 7850.0, 150e6, 1e9, 0.0)
        d.setTemperature(20.0)
        d.setThermal(True, K_T, CP, 0.9)
        if exp_T:
            d.setThermalExpansion(exp_T)
        d.setField("T", T0)
        if cfl:
            d.setCFL(cfl)
        d.init()
    return g, r, T0


def check(g, r, T0, tol):
    dT_g, dT_r = g.field("T") - T0, r.field("T") - T0
    assert np.max(np.abs(dT_r)) > 0.5  # conduction and plastic heat acted
    for name, a, b in [("T increment", dT_g, dT_r), ("u", g.field("u"), r.field("u")),
                       ("pl_strain", g.field("pl_strain"), r.field("pl_strain"))]:
        e = rel(a, b)
        assert e <= tol, f"{name}: {e:.3e}"
    assert np.max(r.field("pl_strain")) > 0


@pytest.mark.parametrize("mesh", [_hexa16, _cyl12])
def test_per_step_path_fp64(mesh):
    say(f"case: thermal JC compression, per-step API, FP64, {mesh.__name__}")
    g, r, T0 = _pair(mesh)
    g.run(100)
    r.run(100)
    check(g, r, T0, 1e-10)


@pytest.mark.parametrize("mesh", [_hexa16, _cyl12])
@pytest.mark.parametrize("cfl", [0.0, 0.25])
def test_fused_path_fp64(mesh, cfl):
    say(f"case: thermal JC compression, fused, FP64, cfl {cfl}, {mesh.__name__}")
    g, r, T0 = _pair(mesh, cfl=cfl)
    g.runFused(40)
    g.runFused(60)
    r.run(100)
    check(g, r, T0, 1e-10)
    assert g.nonfiniteCount() == 0


def test_fused_fp32_mixed():
    say("case: thermal JC compression, fused FP32-mixed vs FP64 oracle")
    g, r, T0 = _pair(_hexa16, precision=WFX_FP32_MIXED)
    g.runFused(100)
    r.run(100)
    dT = rel(g.field("T") - T0, r.field("T") - T0)
    say(f"  FP32-mixed rel err: T increment {dT:.3e} u {rel(g.field('u'), r.field('u')):.3e}")
    check(g, r, T0, 1e-3)


def test_adjoint_rejects_two_way_thermal():
    # the adjoint covers thermal coupling one way only (tests/test_gpu_adjoint.py):
    # a law that reads temperature, or thermal expansion, is refused
    from weldformfem_amd._lib import WfxError

    g, _, _ = _pair(_hexa16)  # Johnson-Cook
    with pytest.raises(WfxError, match="ignores temperature"):
        g.adjointForward(10, checkpoint_every=5)
    g, _, _ = _pair(_hexa16, exp_T=EXP_T, law="bilinear")
    with pytest.raises(WfxError, match="thermal expansion"):
        g.adjointForward(10, checkpoint_every=5)


EXP_T = 1.2e-5  # steel, 1/K

# The trajectories with expansion use the bilinear law: with Johnson-Cook, an
# element still at rest has a strain rate that is exactly the isotropic
# expansion term, and its deviatoric part (the law's effective rate) is then
# rounding noise of ~1e-30 or exactly zero depending on the order of
# operations; 1 + C ln(rate / eps_0) turns that noise into yield stresses of
# either sign, so the fused path (its own sum order) and the reference part
# there although both are right to rounding.  The per-step kernels keep the
# reference's order and follow it there too.


@pytest.mark.parametrize("tet", [False, True])
def test_thermal_expansion_kernel(tet):
    say(f"case: calcThermalExpansion FP64 on a random state after ThermalCalcs, tet={tet}")
    x, el, n = jittered_box(5, 4, 3, 0.02, tet, seed=19)
    g, r = Domain(WFX_FP64), RefDomain()
    for d in (g, r):
        d.CreateFromMesh(x, el, n)
        d.setDensity(7850.0)
        d.setMaterial(0, 200e9, 0.3, 7850.0, 300e6, 1e9, 0.0)
        d.SetDT(1e-4)
        d.init()
        d.setThermal(True, K_T, CP, 0.9)
        d.setThermalExpansion(EXP_T)
    rng = np.random.default_rng(7)
    T0 = 20.0 + 800.0 * rng.random(g.node_count)
    q = 1e9 * rng.random(g.elem_count)
    sr = rng.uniform(-50.0, 50.0, 6 * g.elem_count)
    for d in (g, r):
        d.setField("T", T0)
        d.setField("q_plheat", q)
        d.ThermalCalcs()
        d.setField("str_rate", sr)
        d.calcThermalExpansion()
    gs, rs = g.field("str_rate"), r.field("str_rate")
    assert np.max(np.abs(rs - sr)) > 1e-3  # the rates moved the diagonal
    e = rel(gs - sr, rs - sr)
    assert e <= 1e-13, f"strain-rate change: {e:.3e}"


@pytest.mark.parametrize("mesh", [_hexa16, _cyl12])
@pytest.mark.parametrize("path", ["api", "fused"])
def test_thermal_expansion_trajectory_fp64(mesh, path):
    say(f"case: thermal JC compression with expansion, {path}, FP64, {mesh.__name__}")
    g, r, T0 = _pair(mesh, exp_T=EXP_T, law="bilinear")
    if path == "api":
        g.run(100)
    else:
        g.runFused(30)
        g.runFused(70)
    r.run(100)
    check(g, r, T0, 1e-10)
    e = rel(g.field("eps"), r.field("eps"))
    assert e <= 1e-10, f"eps: {e:.3e}"
    # The shift is isotropic, so the deviatoric update (stress, plastic flow,
    # motion) does not see it: it shows in the strain field.  (Hexahedra; on
    # the tet cylinder the rates it reads, those of its first elements at
    # node ids, sit in the cold bottom layers.)
    if mesh is _hexa16:
        g0, _, _ = _pair(mesh, law="bilinear")
        g0.run(100) if path == "api" else g0.runFused(100)
        assert rel(g0.field("eps"), r.field("eps")) > 1e-6


def test_thermal_expansion_johnson_cook_per_step_path():
    say("case: thermal JC compression with expansion, per-step API (reference order), FP64")
    g, r, T0 = _pair(_hexa16, exp_T=EXP_T)
    g.run(100)
    r.run(100)
    check(g, r, T0, 1e-10)


def test_thermal_expansion_fused_fp32_mixed():
    say("case: thermal JC compression with expansion, fused FP32-mixed vs FP64 oracle")
    g, r, T0 = _pair(_hexa16, precision=WFX_FP32_MIXED, exp_T=EXP_T, law="bilinear")
    g.runFused(100)
    r.run(100)
    check(g, r, T0, 1e-3)
