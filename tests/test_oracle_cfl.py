"""Pins of the oracle's per-step CFL time step (Solver_explicit.C:576-600).

dt = cfl * min edge / (sqrt(K / rho[0]) + max |v|) after the prediction and
the Jacobians; the prediction of the step still uses the previous dt.
Checked on rigid translation (closed forms for every step and for the
positions), on a prescribed velocity, and against the edge-length reduction
on a distorted mesh.  No GPU.
"""
import math

import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import workloads

E, NU, RHO = 200e9, 0.3, 7850.0
K = E / (3 * (1 - 2 * NU))
C = math.sqrt(K / RHO)


def block(nx, ny, nz, h, tet=False):
    d = RefDomain()
    d.AddBoxLength([0, 0, 0], [nx * h, ny * h, nz * h], h / 2, True, tet)
    d.setDensity(RHO)
    d.setMaterial(0, E, NU, RHO, 300e6, 1e9, 0.0)
    return d


@pytest.mark.parametrize("tet", [False, True])
def test_rigid_translation_closed_form(tet):
    h, cfl, dt0 = 0.1, 0.3, 1.0e-6
    d = block(3, 2, 2, h, tet)
    d.SetDT(dt0)
    d.setCFL(cfl)
    d.init()
    v0 = np.array([3.0, 4.0, 0.0])  # |v| = 5: not the largest component, not the sum
    N = d.node_count
    d.setField("v", np.tile(v0, N))
    x0 = d.field("x").reshape(-1, 3).copy()
    want = cfl * h / (C + 5.0)
    moved = np.zeros(3)
    prev = dt0
    for k in range(4):
        d.run(1)
        assert d.getDT() == pytest.approx(want, rel=1e-14)
        # the prediction of step k ran with the dt of step k-1
        moved += prev * v0
        prev = d.getDT()
        np.testing.assert_allclose(d.field("x").reshape(-1, 3), x0 + moved, rtol=0, atol=1e-15)
        # roundoff strain of the moved mesh leaves ~1e-12 accelerations
        np.testing.assert_allclose(d.field("v").reshape(-1, 3), np.tile(v0, (N, 1)), rtol=0, atol=1e-9)


def test_prescribed_velocity_sets_max_speed():
    h, cfl = 0.05, 0.3
    d = block(2, 2, 2, h)
    d.SetDT(1e-7)
    d.setCFL(cfl)
    bottom, top = workloads.box_layers(2, 2, 2)
    for n in bottom:
        d.AddBCVelNode(int(n), 2, 0.0)
    for n in top:
        d.AddBCVelNode(int(n), 2, -10.0)
    d.init()
    d.run(1)
    assert d.getDT() == pytest.approx(cfl * h / (C + 10.0), rel=1e-14)


def test_sound_speed_from_element_zero_density():
    # c = sqrt(K / rho[0]): element 0's density, not the material's
    d = block(2, 1, 1, 0.1)
    d.SetDT(1e-6)
    d.setCFL(0.5)
    d.init()
    rho = d.field("rho")
    rho[0] = 4.0 * RHO
    d.setField("rho", rho)
    d.run(1)
    assert d.getDT() == pytest.approx(0.5 * 0.1 / (math.sqrt(K / (4.0 * RHO)) + 0.0), rel=1e-14)


def test_distorted_mesh_uses_min_edge_reduction():
    r0 = block(4, 3, 3, 0.02, tet=True)
    x = r0.field("x").reshape(-1, 3)
    x = x + (2 * workloads.uniform(5, x.size).reshape(-1, 3) - 1) * 0.2 * 0.02
    d = RefDomain()
    d.CreateFromMesh(x, r0.field("elnod"), 4)
    d.setDensity(RHO)
    d.setMaterial(0, E, NU, RHO, 300e6, 1e9, 0.0)
    d.SetDT(1e-8)
    d.setCFL(0.3)
    d.init()
    minl = d.calcMinEdgeLength()
    assert minl < 0.02
    # brute force over the six edges of every tet
    el = d.field("elnod").reshape(-1, 4)
    P = x[el]
    edges = [P[:, i] - P[:, j] for i in range(4) for j in range(i)]
    assert minl == pytest.approx(min(np.linalg.norm(e, axis=1).min() for e in edges), rel=1e-15)
    d.run(1)
    assert d.getDT() == pytest.approx(0.3 * minl / C, rel=1e-14)


def test_fixed_dt_when_factor_is_zero():
    d = block(2, 1, 1, 0.1)
    d.SetDT(3e-7)
    d.init()
    d.setField("v", np.tile([1.0, 0.0, 0.0], d.node_count))
    d.run(3)
    assert d.getDT() == 3e-7
