"""Pins of the oracle's rate- and temperature-dependent hardening laws.

Johnson-Cook and GMT yield stress and tangent modulus
(Material.cuh:378-473) and the effective strain rate and strain-rate cap of
CalcStressStrain (Mechanical.C:1700-1737), checked against hand-computed
values, against the derivative of the yield law (the tangent modulus is its
derivative wherever no clamp is active) and against the radial-return
invariant.  No GPU.
"""
import math

import numpy as np
import pytest

from oracle.wfref import GMT, JOHNSON_COOK, RefDomain

E, NU, RHO = 200e9, 0.3, 7850.0
G = E / (2 * (1 + NU))
# A, B, n, C, eps_0, m, T_m, T_t
JC = (300e6, 500e6, 0.5, 0.01, 1.0, 1.0, 1520.0, 20.0)
# n1, n2, C1, C2, m1, m2, I1, I2 and the strain, rate, temperature ranges
GMT_C = (-2e-4, 0.25, 150e6, -1.5e-3, 1e-4, 0.02, -1e-5, 0.004)
GMT_R = (0.02, 2.0, 1e-3, 1e4, 20.0, 1200.0)


def one(model, **kw):
    d = RefDomain()
    d.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.05)
    d.setDensity(RHO)
    if model == JOHNSON_COOK:
        d.setMaterialJC(E, NU, RHO, *kw.get("jc", JC))
    else:
        d.setMaterialGMT(E, NU, RHO, 100e6, kw.get("c", GMT_C), kw.get("r", GMT_R))
    d.setTemperature(kw.get("T", 20.0))
    return d


def dev6(D):
    M = np.array([[D[0], D[3], D[5]], [D[3], D[1], D[4]], [D[5], D[4], D[2]]])
    return M - np.trace(M) / 3 * np.eye(3)


def test_effective_strain_rate_is_von_mises_of_D():
    assert RefDomain.effStrainRate([1.0, -0.5, -0.5, 0, 0, 0]) == pytest.approx(1.5, rel=1e-15)
    assert RefDomain.effStrainRate([2.0, 2.0, 2.0, 0, 0, 0]) == 0.0
    assert RefDomain.effStrainRate([0, 0, 0, 0.5, 0, 0]) == pytest.approx(math.sqrt(0.75), rel=1e-15)
    rng = np.random.default_rng(2)
    for _ in range(5):
        D = rng.uniform(-3, 3, 6)
        s = dev6(D)
        assert RefDomain.effStrainRate(D) == pytest.approx(math.sqrt(1.5 * np.sum(s * s)), rel=1e-13)


def test_johnson_cook_yield_hand_values():
    d = one(JOHNSON_COOK)
    # T = T_t: no softening; ep = 0.04 -> A + B*0.2 = 400 MPa; edot = e -> 1 + C
    assert d.yieldStress(0, 0.04, math.e) == pytest.approx(400e6 * 1.01, rel=1e-14)
    # zero rate is read as 1e-5 (Material.cuh:381-383)
    assert d.yieldStress(0, 0.04, 0.0) == pytest.approx(400e6 * (1 + 0.01 * math.log(1e-5)), rel=1e-14)
    # homologous temperature 1/2 with m = 1 halves the stress
    d.setTemperature(770.0)
    assert d.yieldStress(0, 0.04, 1.0) == pytest.approx(200e6, rel=1e-14)
    # A is the element's initial yield
    d.setElemSy0([100e6])
    assert d.yieldStress(0, 0.04, 1.0) == pytest.approx(100e6, rel=1e-14)


@pytest.mark.parametrize("T", [20.0, 400.0])
@pytest.mark.parametrize("ep", [0.003, 0.05, 0.4])
def test_johnson_cook_tangent_is_derivative(T, ep):
    d = one(JOHNSON_COOK, T=T)
    edot = 35.0
    h = 1e-6 * ep
    fd = (d.yieldStress(0, ep + h, edot) - d.yieldStress(0, ep - h, edot)) / (2 * h)
    assert d.tangentModulus(0, ep, edot) == pytest.approx(fd, rel=1e-7)
    # at zero plastic strain the reference takes 0.1 E (Material.cuh:407-408)
    assert d.tangentModulus(0, 0.0, edot) == pytest.approx(0.1 * E, rel=1e-15)


def test_gmt_yield_hand_values():
    # T = 0 inside the range, n1 = m1 = I1 = C2 = 0: C1 * e^n2 * exp(I2/e) * er^m2
    c = (0.0, 0.5, 100e6, 0.0, 0.0, 0.0, 0.0, 0.0)
    d = one(GMT, c=c, r=(0.0, 10.0, 0.0, 1e10, -100.0, 100.0), T=0.0)
    assert d.yieldStress(0, 0.25, 7.0) == pytest.approx(50e6, rel=1e-14)
    # exp(C2 T) with C2 = ln2/100 doubles it at T = 100
    c = (0.0, 0.5, 100e6, math.log(2) / 100, 0.0, 0.0, 0.0, 0.0)
    d = one(GMT, c=c, r=(0.0, 10.0, 0.0, 1e10, -100.0, 100.0), T=100.0)
    assert d.yieldStress(0, 0.25, 7.0) == pytest.approx(100e6, rel=1e-14)
    # rate exponent m2 = 1, I2 = ln(3)*e so exp(I2/e) = 3 at e = 0.25
    c = (0.0, 0.0, 100e6, 0.0, 0.0, 1.0, 0.0, 0.25 * math.log(3))
    d = one(GMT, c=c, r=(0.0, 10.0, 0.0, 1e10, -100.0, 100.0), T=0.0)
    assert d.yieldStress(0, 0.25, 2.0) == pytest.approx(600e6, rel=1e-14)


@pytest.mark.parametrize("T", [20.0, 300.0, 900.0])
@pytest.mark.parametrize("ep", [0.05, 0.3, 1.2])
def test_gmt_tangent_is_derivative(T, ep):
    d = one(GMT, T=T)
    edot = 12.0
    h = 1e-6 * ep
    fd = (d.yieldStress(0, ep + h, edot) - d.yieldStress(0, ep - h, edot)) / (2 * h)
    assert d.tangentModulus(0, ep, edot) == pytest.approx(fd, rel=1e-6)


def test_gmt_clamps():
    d = one(GMT, T=600.0)
    e_min, e_max, r_min, r_max, T_min, T_max = GMT_R
    assert d.yieldStress(0, 0.0, 5.0) == d.yieldStress(0, e_min, 5.0)
    assert d.yieldStress(0, 9.0, 5.0) == d.yieldStress(0, e_max, 5.0)
    assert d.yieldStress(0, 0.1, 0.0) == d.yieldStress(0, 0.1, r_min)
    assert d.yieldStress(0, 0.1, 1e9) == d.yieldStress(0, 0.1, r_max)
    assert d.tangentModulus(0, 9.0, 5.0) == d.tangentModulus(0, e_max, 5.0)
    hot = d.yieldStress(0, 0.1, 5.0)
    d.setTemperature(5000.0)
    assert d.yieldStress(0, 0.1, 5.0) != hot
    d2 = one(GMT, T=T_max)
    assert d.yieldStress(0, 0.1, 5.0) == d2.yieldStress(0, 0.1, 5.0)
    d.setTemperature(-50.0)
    d2.setTemperature(T_min)
    assert d.yieldStress(0, 0.1, 5.0) == d2.yieldStress(0, 0.1, 5.0)


def von_mises(t6):
    return math.sqrt(1.5 * np.sum(dev6(t6) ** 2))


@pytest.mark.parametrize("model", [JOHNSON_COOK, GMT])
@pytest.mark.parametrize("cap", [1e6, 100.0])
def test_radial_return_with_rate_dependent_yield(model, cap):
    # the yield stress takes the uncapped rate, the tangent the capped one
    # (Mechanical.C:1725-1748); after the return the equivalent stress sits on
    # sigma_y + H dgamma
    d = one(model, T=250.0)
    d.setMaxStrainRate(cap)
    D = 60.0 * np.array([-40.0, 15.0, 25.0, 30.0, -5.0, 12.0])
    edot = RefDomain.effStrainRate(D)
    assert edot > 100.0
    ep0, dt = 0.08, 4e-6
    d.setField("str_rate", D)
    d.setField("rot_rate", np.zeros(6))
    d.setField("tau", np.zeros(6))
    d.setField("p", [0.0])
    d.setField("pl_strain", [ep0])
    d.CalcStressStrain(dt)
    tau, ep = d.field("tau"), d.field("pl_strain")[0]
    sy = d.yieldStress(0, ep0, edot)
    H = d.tangentModulus(0, ep0, min(edot, cap))
    assert d.field("sigma_y")[0] == pytest.approx(sy, rel=1e-15)
    s_tr = von_mises(2 * G * dt * D)
    assert s_tr > sy
    dg = (s_tr - sy) / (3 * G + H)
    assert ep - ep0 == pytest.approx(dg, rel=1e-12)
    assert von_mises(tau) == pytest.approx(sy + H * dg, rel=1e-11)
    if cap < edot:
        assert H != pytest.approx(d.tangentModulus(0, ep0, edot), rel=1e-6)
