"""Pins of the oracle's thermal coupling (Thermal.C:27-115,
Mechanical.C:1784-1810).

Conduction: a uniform field stays put, sum m T is conserved, and a linear
field along a bar gives zero rate at interior nodes and the Fourier flux
k g A into the cold end.  Plastic heat: frac * dep * sigma_vm^2 / (sigma_y
dt), which for perfect plasticity is the textbook frac * sigma_y * dep / dt.
Heating: 0.9 q dt / (rho cp) on a lone element.  Law temperature: the mean of
the element's nodal temperatures.  No GPU.
"""
import math

import numpy as np
import pytest

from oracle.wfref import JOHNSON_COOK, RefDomain

E_, NU, RHO = 200e9, 0.3, 7850.0
G = E_ / (2 * (1 + NU))
K_T, CP = 50.0, 460.0


def block(nx, ny, nz, h, tet=False):
    d = RefDomain()
    d.AddBoxLength([0, 0, 0], [nx * h, ny * h, nz * h], h / 2, True, tet)
    d.setDensity(RHO)
    d.setMaterial(0, E_, NU, RHO, 300e6, 0.0, 0.0)
    d.SetDT(1e-3)
    d.init()
    d.setThermal(True, K_T, CP, 0.9)
    return d


@pytest.mark.parametrize("tet", [False, True])
def test_uniform_field_and_conservation(tet):
    d = block(3, 2, 2, 0.1, tet)
    N = d.node_count
    d.setField("T", np.full(N, 300.0))
    d.ThermalCalcs()
    np.testing.assert_allclose(d.field("T"), 300.0, rtol=1e-12)
    T0 = 20.0 + 500.0 * np.random.default_rng(3).random(N)
    d.setField("T", T0)
    m = d.field("mdiag")
    d.ThermalCalcs()
    T1 = d.field("T")
    assert np.max(np.abs(T1 - T0)) > 1e-3  # conduction acts
    assert np.sum(m * T1) == pytest.approx(np.sum(m * T0), rel=1e-13)


def test_linear_field_fourier_flux():
    h, nx, ny, nz, g = 0.05, 4, 2, 1, 1000.0
    d = block(nx, ny, nz, h)
    x = d.field("x").reshape(-1, 3)
    T0 = 100.0 + g * x[:, 0]
    d.setField("T", T0)
    dt = 1e-3
    m = d.field("mdiag")
    d.ThermalCalcs()
    rate = m * (d.field("T") - T0) / dt
    L = nx * h
    cold, hot = np.isclose(x[:, 0], 0.0), np.isclose(x[:, 0], L)
    interior = ~(cold | hot)
    area = (ny * h) * (nz * h)
    assert np.max(np.abs(rate[interior])) < 1e-9 * K_T * g * area
    assert np.sum(rate[cold]) == pytest.approx(K_T * g * area, rel=1e-11)
    assert np.sum(rate[hot]) == pytest.approx(-K_T * g * area, rel=1e-11)


def von_mises(t6):
    s = np.array(t6, float)
    dev = s.copy()
    dev[:3] -= s[:3].mean()
    return math.sqrt(1.5 * (dev[0] ** 2 + dev[1] ** 2 + dev[2] ** 2 + 2 * (dev[3] ** 2 + dev[4] ** 2 + dev[5] ** 2)))


@pytest.mark.parametrize("Et", [0.0, 2e9])
def test_plastic_heat_source(Et):
    d = RefDomain()
    d.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.05)
    d.setDensity(RHO)
    d.setMaterial(0, E_, NU, RHO, 300e6, Et, 0.0)
    d.setThermal(True, K_T, CP, 0.85)
    D = 20.0 * np.array([-40.0, 15.0, 25.0, 30.0, -5.0, 12.0])
    dt = 4e-6
    d.setField("str_rate", D)
    d.setField("rot_rate", np.zeros(6))
    d.setField("tau", np.zeros(6))
    d.setField("p", [-1e8])
    d.setField("pl_strain", [0.01])
    d.CalcStressStrain(dt)
    dep = d.field("pl_strain")[0] - 0.01
    assert dep > 0
    sy = d.field("sigma_y")[0]
    q = d.field("q_plheat")[0]
    assert q == pytest.approx(0.85 * dep * von_mises(d.field("sigma")) ** 2 / (sy * dt), rel=1e-11)
    if Et == 0.0:
        assert q == pytest.approx(0.85 * 300e6 * dep / dt, rel=1e-11)
    # elastic step: no heat
    d.setField("str_rate", 1e-3 * D)
    d.setField("tau", np.zeros(6))
    d.CalcStressStrain(dt)
    assert d.field("q_plheat")[0] == 0.0


def test_heating_of_a_lone_element():
    d = block(1, 1, 1, 0.1)
    d.setThermal(True, 0.0, CP, 0.9)
    d.setField("T", np.full(8, 20.0))
    q = 3.0e9
    d.setField("q_plheat", [q])
    d.ThermalCalcs()
    np.testing.assert_allclose(d.field("T"), 20.0 + 0.9 * q * 1e-3 / (RHO * CP), rtol=1e-12)


def test_law_reads_mean_nodal_temperature():
    d = RefDomain()
    d.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.05)
    d.setDensity(RHO)
    d.setMaterialJC(E_, NU, RHO, 300e6, 500e6, 0.5, 0.0, 1.0, 1.0, 1520.0, 20.0)
    d.setTemperature(20.0)
    d.setThermal(True, K_T, CP, 0.9)
    d.setField("T", [20.0, 20.0, 20.0, 20.0, 1520.0, 1520.0, 1520.0, 1520.0])  # mean 770: T_h = 1/2
    assert d.yieldStress(0, 0.04, 1.0) == pytest.approx(200e6, rel=1e-14)
    d.setThermal(False, K_T, CP, 0.9)
    assert d.yieldStress(0, 0.04, 1.0) == pytest.approx(400e6, rel=1e-14)


# ---- calcThermalExpansion (Thermal.C:164-178) --------------------------------
# The strain rate loses exp_T times the mean of the element-node rates read at
# m_dTedt[node], the entries of the element-node array whose flat index is
# one of the element's node ids (the array is laid out [element][local node]).

def _expansion_state(tet, seed):
    d = block(3, 2, 2, 0.1, tet)
    rng = np.random.default_rng(seed)
    n, E = (4 if tet else 8), d.elem_count
    sr = rng.uniform(-50.0, 50.0, 6 * E)
    rates = rng.uniform(-10.0, 10.0, n * E)
    d.setField("str_rate", sr)
    d.setField("dTedt", rates)
    return d, sr, rates, n, E


@pytest.mark.parametrize("tet", [False, True])
def test_expansion_zero_rates_and_zero_coefficient(tet):
    d, sr, rates, n, E = _expansion_state(tet, 4)
    d.setThermalExpansion(0.0)
    d.calcThermalExpansion()
    np.testing.assert_array_equal(d.field("str_rate"), sr)
    d.setThermalExpansion(1.2e-5)
    d.setField("dTedt", np.zeros(n * E))
    d.calcThermalExpansion()
    np.testing.assert_array_equal(d.field("str_rate"), sr)


@pytest.mark.parametrize("tet", [False, True])
def test_expansion_uniform_rate_is_isotropic(tet):
    # every rate r: each element's diagonal strain rates drop by exp_T r, the
    # shear rates stay; the volumetric rate drops by 3 exp_T r (free thermal
    # expansion removed from the mechanical strain rate)
    d, sr, rates, n, E = _expansion_state(tet, 5)
    a, r = 1.2e-5, 7.5
    d.setThermalExpansion(a)
    d.setField("dTedt", np.full(n * E, r))
    d.calcThermalExpansion()
    got = d.field("str_rate").reshape(E, 6)
    want = sr.reshape(E, 6).copy()
    want[:, :3] -= a * r
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(got[:, 3:], sr.reshape(E, 6)[:, 3:])
    vol = got[:, :3].sum(axis=1) - sr.reshape(E, 6)[:, :3].sum(axis=1)
    np.testing.assert_allclose(vol, -3 * a * r, rtol=1e-9)


@pytest.mark.parametrize("tet", [False, True])
def test_expansion_reads_the_entries_at_the_node_ids(tet):
    # distinct rates: an element sees exactly the entries whose flat index is
    # one of its node ids, not its own element-node entries
    d, sr, rates, n, E = _expansion_state(tet, 6)
    a = 2.0e-5
    d.setThermalExpansion(a)
    d.calcThermalExpansion()
    el = d.field("elnod").reshape(E, n)
    mean_at_nodes = rates[el].mean(axis=1)
    own = rates.reshape(E, n).mean(axis=1)
    got = d.field("str_rate").reshape(E, 6)[:, 0]
    np.testing.assert_allclose(got, sr.reshape(E, 6)[:, 0] - a * mean_at_nodes, rtol=0, atol=1e-12)
    assert np.max(np.abs(got - (sr.reshape(E, 6)[:, 0] - a * own))) > 1e-6


def test_expansion_runs_in_the_step_after_the_strain_rates():
    # a heated block (plastic heat) with expansion: the step's stress update
    # sees the reduced strain rate, so the run differs from exp_T = 0
    def run(a):
        d = RefDomain()
        d.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.0125)
        from weldformfem_amd import workloads
        bottom, top = workloads.box_layers(4, 4, 4)
        workloads.setup_compression(d, bottom, top, -20.0, sy0=None, hg_coeff=0.0)
        d.init()
        d.setThermal(True, K_T, CP, 0.9)
        d.setThermalExpansion(a)
        d.run(60)
        return d.field("u"), d.field("T")
    u0, T0 = run(0.0)
    u1, T1 = run(1.2e-5)
    assert np.max(T0) > 20.0  # plastic heat raised the temperature
    assert np.max(np.abs(u1 - u0)) > 0
