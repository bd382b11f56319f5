"""Hollomon, Johnson-Cook and GMT hardening on the GPU against the oracle.

The stress update alone on a random state (per-step API), then 100-step
trajectories through the per-step API and the fused path, FP64 (1e-10
relative, as for the bilinear law; the device pow/exp/log may differ from
glibc in the last bit) and FP32-mixed (1e-4 displacement, 1e-3 reaction and
plastic strain).  Per-element temperatures and the strain-rate cap of the
tangent are exercised.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, WFX_HOLLOMON, workloads

pytestmark = pytest.mark.gpu

E_, NU, RHO = 200e9, 0.3, 7850.0
JC = (300e6, 500e6, 0.3, 0.014, 1.0, 1.0, 1520.0, 20.0)  # A B n C eps_0 m T_m T_t
GMT_C = (-2e-4, 0.15, 600e6, -1.5e-3, 1e-4, 0.01, -1e-5, 1e-4)  # n1 n2 C1 C2 m1 m2 I1 I2
GMT_R = (0.002, 3.0, 1e-3, 1e5, 20.0, 1200.0)


def say(msg):
    print(msg, flush=True)


def set_law(d, law):
    if law == "hollomon":
        d.setMaterial(WFX_HOLLOMON, E_, NU, RHO, 300e6, 900e6, 0.2)
    elif law == "jc":
        d.setMaterialJC(E_, NU, RHO, *JC)
    else:
        d.setMaterialGMT(E_, NU, RHO, 250e6, GMT_C, GMT_R)


def temperatures(law, E):
    if law == "hollomon":
        return None
    return 20.0 + 580.0 * workloads.uniform(21, E)


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


@pytest.mark.parametrize("law", ["hollomon", "jc", "gmt"])
def test_stress_update_random_state(law):
    say(f"case: stress update FP64 on a random state, law {law}")
    r0 = RefDomain()
    r0.AddBoxLength([0, 0, 0], [0.1, 0.08, 0.06], 0.01)
    x, el = r0.field("x"), r0.field("elnod")
    g, r = Domain(WFX_FP64), RefDomain()
    for d in (g, r):
        d.CreateFromMesh(x, el, 8)
        d.setDensity(RHO)
        set_law(d, law)
    E = g.elem_count
    T = temperatures(law, E)
    if T is not None:
        for d in (g, r):
            d.setTemperature(T)
            d.setMaxStrainRate(2000.0)
    rng = np.random.default_rng(6)
    vals = {"tau": rng.uniform(-2e8, 2e8, 6 * E), "eps": rng.uniform(-1e-2, 1e-2, 6 * E),
            "p": rng.uniform(-2e8, 2e8, E), "pl_strain": rng.uniform(0, 0.2, E),
            "str_rate": rng.uniform(-3000, 3000, 6 * E)}
    rot = np.zeros(6 * E)
    rot.reshape(-1, 6)[:, 3:] = rng.uniform(-50, 50, (E, 3))
    vals["rot_rate"] = rot
    vals["pl_strain"][:5] = 0.0  # Johnson-Cook's zero-strain tangent branch
    for k, v in vals.items():
        g.setField(k, v)
        r.setField(k, v)
    for d in (g, r):
        d.CalcStressStrain(2.0e-6)
    assert np.any(r.field("pl_strain") > vals["pl_strain"])
    for f in ["sigma_y", "pl_strain", "tau", "sigma", "strain_pl_incr", "eps"]:
        e = rel(g.field(f), r.field(f))
        assert e <= 1e-13, f"{f}: {e:.3e}"


def _hexa20():
    r = RefDomain()
    r.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], 0.025)
    bottom, top = workloads.box_layers(20, 20, 20)
    return r.field("x"), r.field("elnod"), 8, bottom, top


def _cyl16():
    x, el, bottom, top = workloads.cylinder_tets(cells=16)
    return x, el, 4, bottom, top


def _pair(mesh, law, precision, cap=None):
    x, el, n, bottom, top = mesh()
    g = Domain(precision)
    g.CreateFromMesh(x, el, n)
    dt = workloads.setup_compression(g, bottom, top, -20.0)
    r = RefDomain()
    r.CreateFromMesh(x, el, n)
    workloads.setup_compression(r, bottom, top, -20.0, dt=dt)
    T = temperatures(law, g.elem_count)
    for d in (g, r):
        set_law(d, law)
        if T is not None:
            d.setTemperature(T)
        if cap is not None:
            d.setMaxStrainRate(cap)
        d.init()
    return g, r


@pytest.mark.parametrize("law", ["hollomon", "jc", "gmt"])
@pytest.mark.parametrize("mesh", [_hexa20, _cyl16])
def test_trajectory_100_steps_fp64(law, mesh):
    say(f"case: per-step API 100-step FP64 trajectory, law {law}, {mesh.__name__}")
    g, r = _pair(mesh, law, WFX_FP64)
    rg, rr = g.run(100), r.run(100)
    assert np.max(r.field("pl_strain")) > 0
    for f in ["u", "v", "pl_strain", "sigma_y"]:
        e = rel(g.field(f), r.field(f))
        assert e <= 1e-10, f"{f}: {e:.3e}"
    assert rel(rg, rr) <= 1e-10


@pytest.mark.parametrize("law,cap", [("hollomon", None), ("jc", None), ("jc", 5.0), ("gmt", None),
                                     ("gmt", 5.0)])
@pytest.mark.parametrize("mesh", [_hexa20, _cyl16])
def test_fused_trajectory_100_steps_fp64(law, cap, mesh):
    say(f"case: fused 100-step FP64 trajectory, law {law}, cap {cap}, {mesh.__name__}")
    g, r = _pair(mesh, law, WFX_FP64, cap=cap)
    rg = np.concatenate([g.runFused(30), g.runFused(70)])
    rr = r.run(100)
    assert np.max(r.field("pl_strain")) > 0
    for f in ["u", "v", "x", "pl_strain", "sigma_y", "tau", "p"]:
        e = rel(g.field(f), r.field(f))
        assert e <= 1e-10, f"{f}: {e:.3e}"
    assert rel(rg, rr) <= 1e-10
    assert g.nonfiniteCount() == 0


@pytest.mark.parametrize("law", ["hollomon", "jc", "gmt"])
def test_fused_trajectory_fp32_mixed(law):
    say(f"case: fused 100-step FP32-mixed vs FP64 oracle, law {law}")
    g, r = _pair(_hexa20, law, WFX_FP32_MIXED)
    rg, rr = g.runFused(100), r.run(100)
    eu, er, ep = rel(g.field("u"), r.field("u")), rel(rg, rr), rel(g.field("pl_strain"), r.field("pl_strain"))
    say(f"  FP32-mixed rel err: u {eu:.3e} reaction {er:.3e} pl_strain {ep:.3e}")
    assert eu <= 1e-4 and er <= 1e-3 and ep <= 1e-3


def test_setters_reject_bad_input():
    from weldformfem_amd._lib import WfxError

    g = Domain(WFX_FP64)
    g.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.05)
    with pytest.raises(WfxError, match="eps_0"):
        g.setMaterialJC(E_, NU, RHO, 300e6, 500e6, 0.3, 0.01, 0.0, 1.0, 1520.0, 20.0)
    with pytest.raises(WfxError, match="one value per element"):
        g.setTemperature(np.ones(3))
    with pytest.raises(WfxError, match="non-negative"):
        g.setMaxStrainRate(-1.0)
    with pytest.raises(ValueError):
        g.setMaterialGMT(E_, NU, RHO, 250e6, GMT_C[:7], GMT_R)
