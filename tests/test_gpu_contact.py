"""Rigid-surface penalty contact on the GPU against the oracle.

The contact step alone on a prepared state (FP64, relative 1e-14), then a
tetrahedral cylinder compressed by a descending rigid plane with friction:
per-step API and fused path over 150 steps (15 refreshes of the external
face areas, runs split off the 10-step rhythm), with the per-step CFL time
step, FP32-mixed, and captured graphs.  FP64 bar 1e-10 relative on u, v,
the contact forces and the reaction; FP32-mixed 1e-3.  Contact heat
conduction (a hot die heating the nodes it touches) on both paths: the
nodal conduction terms and temperatures at 1e-10 against the oracle.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads

pytestmark = pytest.mark.gpu

E_, NU, RHO = 200e9, 0.3, 7850.0


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def tilted_plane(z0, slope=0.05, lo=-2.0, hi=2.0):
    """two triangles facing down (normal about -z), slightly tilted"""
    nodes = np.array([[lo, lo, z0 + slope * lo], [hi, lo, z0 + slope * lo], [lo, hi, z0 + slope * lo],
                      [hi, hi, z0 + slope * lo]])
    nodes[:, 2] = z0 + slope * nodes[:, 0]
    return nodes, np.array([[0, 2, 1], [1, 2, 3]])


def test_contact_step_on_a_prepared_state():
    say("case: contact kernel FP64 on a prepared state (jittered tet box, tilted plane)")
    ref = RefDomain()
    ref.AddBoxLength([0, 0, 0], [0.2, 0.16, 0.12], 0.01, True, True)
    x = ref.field("x").reshape(-1, 3)
    x = x + (2.0 * workloads.uniform(31, x.size).reshape(-1, 3) - 1.0) * 0.15 * 0.02
    el = ref.field("elnod")
    g, r = Domain(WFX_FP64), RefDomain()
    nodes, tris = tilted_plane(0.118)
    for d in (g, r):
        d.CreateFromMesh(x, el, 4)
        d.setDensity(RHO)
        d.setMaterial(0, E_, NU, RHO, 300e6, 1e9, 0.0)
        d.SetDT(2e-7)
        d.setRigidMesh(nodes, tris, vel=[0.0, 0.0, -1.0])
        d.setContactParams(0.1, 0.3, 0.2)
        d.init()
    rng = np.random.default_rng(8)
    for it in range(3):
        v = rng.uniform(-30, 30, 3 * g.node_count)
        for d in (g, r):
            d.setField("v", v)
            d.CalcExtFaceAreas()
            d.CalcContactForces()
        assert np.sum(r.field("mesh_in_contact") == 0) > 10
        np.testing.assert_array_equal(g.field("mesh_in_contact"), r.field("mesh_in_contact"))
        for f in ["node_area", "elem_length", "contforce", "ut_prev", "cfn"]:
            e = rel(g.field(f), r.field(f))
            assert e <= 1e-14, f"call {it} {f}: {e:.3e}"
        assert g.contactForce(0) == pytest.approx(r.contactForce(0), rel=1e-13)
        for d in (g, r):
            d.rigidUpdate()
        np.testing.assert_allclose(g.field("rigid_x"), r.field("rigid_x"), rtol=1e-15)
        np.testing.assert_allclose(g.field("rigid_pplane"), r.field("rigid_pplane"), rtol=1e-14)


def _pair(precision=WFX_FP64, cfl=0.0, cells=12, stream=0):
    x, el, bottom, top = workloads.cylinder_tets(cells=cells)
    zt = np.asarray(x).reshape(-1, 3)[:, 2].max()
    nodes = np.array([[-1, -1, zt], [1, -1, zt], [-1, 1, zt], [1, 1, zt]], float)
    tris = np.array([[0, 2, 1], [1, 2, 3]])
    g = Domain(precision, stream=stream)
    r = RefDomain()
    for d in (g, r):
        d.CreateFromMesh(x, el, 4)
        d.setDensity(RHO)
        d.setMaterial(0, E_, NU, RHO, 300e6, 1e9, 0.0)
        for nd in bottom:
            d.AddBCVelNode(int(nd), 2, 0.0)
        if hasattr(d, "AllocateBCs"):
            d.AllocateBCs()
        d.setReactionNodes(bottom)
    dt = workloads.cfl_dt(r.calcMinEdgeLength(), E_, NU, RHO, 4)
    for d in (g, r):
        d.SetDT(dt)
        d.setRigidMesh(nodes, tris, vel=[0.0, 0.0, -20.0])
        d.setContactParams(0.1, 0.2, 0.1)
        if cfl:
            d.setCFL(cfl)
        d.init()
    return g, r


def check(g, r, hg, hr, tol):
    assert np.sum(r.field("mesh_in_contact") == 0) > 50
    for f in ["u", "v", "contforce", "pl_strain"]:
        e = rel(g.field(f), r.field(f))
        assert e <= tol, f"{f}: {e:.3e}"
    assert rel(hg, hr) <= tol
    assert g.contactForce(0) == pytest.approx(r.contactForce(0), rel=max(tol, 1e-13))


def test_per_step_path_fp64():
    say("case: contact compression, per-step API, FP64, 150 steps")
    g, r = _pair()
    hg, hr = g.run(150), r.run(150)
    check(g, r, hg, hr, 1e-10)


@pytest.mark.parametrize("cfl", [0.0, 0.25])
def test_fused_path_fp64(cfl):
    say(f"case: contact compression, fused FP64, cfl {cfl}, runs of 7, 53, 90 steps")
    g, r = _pair(cfl=cfl)
    hg = np.concatenate([g.runFused(7), g.runFused(53), g.runFused(90)])
    hr = r.run(150)
    check(g, r, hg, hr, 1e-10)
    assert g.nonfiniteCount() == 0


def test_fused_fp32_mixed():
    say("case: contact compression, fused FP32-mixed vs FP64 oracle")
    g, r = _pair(precision=WFX_FP32_MIXED)
    hg, hr = g.runFused(150), r.run(150)
    say(f"  FP32-mixed rel err: u {rel(g.field('u'), r.field('u')):.3e} reaction {rel(hg, hr):.3e} "
        f"contact {abs(g.contactForce(0) / r.contactForce(0) - 1):.3e}")
    for f in ["u", "contforce"]:
        e = rel(g.field(f), r.field(f))
        assert e <= 1e-3, f"{f}: {e:.3e}"
    assert rel(hg, hr) <= 1e-3


def test_graph_replays_and_step_rhythm():
    say("case: contact compression, captured graph of 20 steps, two replays vs eager")
    import torch
    from weldformfem_amd._lib import WfxError

    stream = torch.cuda.Stream()
    g, _ = _pair(stream=stream.cuda_stream)
    e, _ = _pair()
    g.graphCapture(20)
    h = np.concatenate([g.graphLaunch(), g.graphLaunch()])
    torch.cuda.synchronize()
    he = np.concatenate([e.runFused(20), e.runFused(20)])
    np.testing.assert_array_equal(h, he)
    np.testing.assert_array_equal(g.field("contforce"), e.field("contforce"))
    g.graphCapture(15)
    g.graphLaunch()
    with pytest.raises(WfxError, match="mod 10"):
        g.graphLaunch()


def test_hexa_body_rejected():
    from weldformfem_amd._lib import WfxError

    g = Domain(WFX_FP64)
    g.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.025)
    n, t = tilted_plane(0.1)
    with pytest.raises(WfxError, match="tetrahedral"):
        g.setRigidMesh(n, t)


def test_forging_combination_fused_fp64():
    # the features together, as in a forging deck: Johnson-Cook hardening,
    # thermal coupling, per-step CFL time step, a descending die with friction
    say("case: JC + thermal + per-step CFL + contact, fused FP64 vs oracle, 120 steps")
    g, r = _pair(cfl=0.25)
    for d in (g, r):
        d.setMaterialJC(E_, NU, RHO, 300e6, 500e6, 0.3, 0.014, 1.0, 1.0, 1520.0, 20.0)
        d.setTemperature(20.0)
        d.setThermal(True, 5.0e4, 460.0, 0.9)
        d.setCFL(0.25)
        d.init()
    hg = np.concatenate([g.runFused(33), g.runFused(87)])
    hr = r.run(120)
    check(g, r, hg, hr, 1e-10)
    dT = r.field("T") - 20.0
    assert np.max(np.abs(dT)) > 0.01
    assert rel(g.field("T") - 20.0, dT) <= 1e-10
    assert g.getDT() == pytest.approx(r.getDT(), rel=1e-14)


def test_second_surface_replaces_the_first():
    say("case: wfx_set_rigid_mesh twice, the second surface replaces the first")
    g, r = _pair()
    x = np.asarray(workloads.cylinder_tets(cells=12)[0]).reshape(-1, 3)
    zt = x[:, 2].max()
    nodes, tris = tilted_plane(zt, slope=0.02)
    for d in (g, r):
        d.setRigidMesh(nodes, tris, vel=[0.0, 0.0, -15.0])
        d.init()
    hg, hr = g.runFused(60), r.run(60)
    assert np.sum(r.field("mesh_in_contact") == 0) > 10
    for f in ["u", "contforce"]:
        e = rel(g.field(f), r.field(f))
        assert e <= 1e-10, f"{f}: {e:.3e}"
    np.testing.assert_allclose(g.field("rigid_x"), r.field("rigid_x"), rtol=1e-14)


@pytest.mark.parametrize("path", ["api", "fused"])
def test_contact_heat_conduction_fp64(path):
    # a die at 900 C pressed onto a body at 20 C: the nodes it touches take
    # heat_cond * area * (T_die - T) and warm up; conduction inside the body
    # spreads it (Contact.C:309, Thermal.C:124)
    say(f"case: contact heat conduction, {path} FP64 vs oracle, 100 steps")
    g, r = _pair()
    for d in (g, r):
        d.setThermal(True, 50.0, 460.0, 0.9)
        d.setContactHeat(2.0e6, 900.0)
        d.init()
    hg = g.run(100) if path == "api" else np.concatenate([g.runFused(41), g.runFused(59)])
    hr = r.run(100)
    check(g, r, hg, hr, 1e-10)
    q = r.field("q_cont_conv")
    assert np.sum(q > 0) > 50
    e = rel(g.field("q_cont_conv"), q)
    assert e <= 1e-10, f"q_cont_conv: {e:.3e}"
    dT = r.field("T") - 20.0
    assert np.max(dT) > 1e-3
    e = rel(g.field("T") - 20.0, dT)
    assert e <= 1e-10, f"T: {e:.3e}"
