"""Pins of the oracle's rigid-surface penalty contact (Contact.C:31-335,
Domain_d.C:110-296, :2242-2276, Mesh.h:236-330).

External faces and nodal areas against the box surface, element heights
against 3V/A, rigid-plane normals and offsets, and the contact force of one
node: no force above a plane or outside a triangle, the mass-limited and
geometry-limited penalty stiffness, critical damping ratio 0.2, stick and
slip friction, the first-triangle rule and the persistence of a force after
separation; contact heat conduction q = h A (T_die - T) of the nodes in
contact, its place in the nodal temperature update (the heat q dt entering
as m cp dT) and its persistence.  No GPU.
"""
import math

import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import workloads

E_, NU, RHO = 200e9, 0.3, 7850.0


def body(nx=2, ny=2, nz=2, h=0.1, dt=1e-6):
    d = RefDomain()
    d.AddBoxLength([0, 0, 0], [nx * h, ny * h, nz * h], h / 2, True, True)
    d.setDensity(RHO)
    d.setMaterial(0, E_, NU, RHO, 300e6, 1e9, 0.0)
    d.SetDT(dt)
    return d


def plane(z, lo=-1.0, hi=2.0, flip=False):
    """two triangles spanning [lo, hi]^2 at height z, normal +z (or -z)"""
    nodes = [[lo, lo, z], [hi, lo, z], [lo, hi, z], [hi, hi, z]]
    tris = [[0, 1, 2], [1, 3, 2]] if not flip else [[0, 2, 1], [1, 2, 3]]
    return np.array(nodes), np.array(tris)


def test_external_faces_and_areas():
    d = body(3, 2, 2, 0.1)
    n, t = plane(-1.0)
    d.setRigidMesh(n, t)
    x = d.field("x").reshape(-1, 3)
    on_surface = (np.isclose(x, 0.0) | np.isclose(x, [0.3, 0.2, 0.2])).any(axis=1)
    np.testing.assert_array_equal(d.field("ext_node").astype(bool), on_surface)
    area = d.field("node_area")
    assert np.all(area[~on_surface] == 0.0)
    assert area.sum() == pytest.approx(2 * (0.3 * 0.2 + 0.2 * 0.2 + 0.3 * 0.2), rel=1e-13)


def test_element_heights():
    r0 = body(2, 2, 1, 0.05)
    x = r0.field("x").reshape(-1, 3)
    x = x + (2 * workloads.uniform(7, x.size).reshape(-1, 3) - 1) * 0.2 * 0.05
    d = RefDomain()
    d.CreateFromMesh(x, r0.field("elnod"), 4)
    d.calcElemLength()
    el = d.field("elnod").reshape(-1, 4)
    P = x[el]
    vol6 = np.abs(np.einsum("ei,ei->e", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])))
    h = []
    for i in range(4):
        f = [k for k in range(4) if k != i]
        area2 = np.linalg.norm(np.cross(P[:, f[1]] - P[:, f[0]], P[:, f[2]] - P[:, f[0]]), axis=1)
        h.append(vol6 / area2)  # 3V / A with V = vol6/6 and A = area2/2
    np.testing.assert_allclose(d.field("elem_length"), np.min(h, axis=0), rtol=1e-12)


def test_rigid_plane_normal_offset_and_motion():
    d = body()
    n, t = plane(0.35)
    d.setRigidMesh(n, t, vel=[0.0, 0.0, -2.0])
    np.testing.assert_allclose(d.field("rigid_normal").reshape(-1, 3), [[0, 0, 1], [0, 0, 1]], atol=1e-15)
    np.testing.assert_allclose(d.field("rigid_pplane"), [0.35, 0.35], rtol=1e-15)
    d.rigidUpdate()
    np.testing.assert_allclose(d.field("rigid_pplane"), [0.35 - 2e-6] * 2, rtol=1e-14)
    n2, t2 = plane(0.0, flip=True)
    d.setRigidMesh(n2, t2)
    np.testing.assert_allclose(d.field("rigid_normal").reshape(-1, 3), [[0, 0, -1], [0, 0, -1]], atol=1e-15)


def one_node_case(dt, delta=1e-4, v=(0.0, 0.0, 0.0), pf=0.1, mu=(0.0, 0.0), tri=None):
    """top-corner node of a 1x1x1 six-tet cube pushed delta below a +z plane
    facing down (normal -z) placed at the top face"""
    d = body(1, 1, 1, 0.1, dt=dt)
    x = d.field("x").reshape(-1, 3)
    top = np.where(np.isclose(x[:, 2], 0.1))[0]
    k = top[np.argmax(x[top, 0] + x[top, 1])]  # the (0.1, 0.1, 0.1) corner
    n, t = plane(0.1 - delta, flip=True) if tri is None else tri
    d.setRigidMesh(n, t)
    d.setContactParams(pf, *mu)
    d.init()
    vv = np.zeros_like(x)
    vv[k] = v
    d.setField("v", vv.reshape(-1))
    # only node k penetrates: lift the other top nodes above the plane
    xx = x.copy()
    others = [i for i in top if i != k]
    xx[others, 2] = 0.1 - 2 * delta
    xx[k, 2] = 0.1
    d.setField("x", xx.reshape(-1))
    d.CalcContactForces()
    return d, k


def stiffness(d, k, dt):
    m = d.field("mdiag")[k]
    el = d.field("elnod").reshape(-1, 4)
    nodlen = d.field("elem_length")[np.where((el == k).any(axis=1))[0]].mean()
    kgeo = E_ * d.field("node_area")[k] / nodlen
    return m, kgeo, 0.2 * m / dt ** 2


@pytest.mark.parametrize("dt", [1e-7, 1e-3])
def test_static_penetration_force(dt):
    d, k = one_node_case(dt)
    m, kgeo, kmass = stiffness(d, k, dt)
    kc = min(kgeo, kmass)
    assert (kmass < kgeo) == (dt > 1e-5)  # both branches are exercised
    cf = d.field("contforce").reshape(-1, 3)
    # behind the plane by delta along the normal n = -z: d = -delta and the
    # force -(contPF kc d) n pushes the node down, out of the surface
    np.testing.assert_allclose(cf[k], [0, 0, -0.1 * kc * 1e-4], rtol=1e-12, atol=0)
    others = np.delete(np.arange(d.node_count), k)
    assert np.all(cf[others] == 0.0)
    assert d.field("mesh_in_contact")[k] == 0
    assert d.contactForce(0) == pytest.approx(0.1 * kc * 1e-4, rel=1e-12)


def test_damping_is_critical_ratio_02():
    dt = 1e-7
    d0, k = one_node_case(dt)
    d1, _ = one_node_case(dt, v=(0.0, 0.0, 0.5))
    m, kgeo, kmass = stiffness(d0, k, dt)
    kc = min(kgeo, kmass)
    dz = d1.field("contforce").reshape(-1, 3)[k, 2] - d0.field("contforce").reshape(-1, 3)[k, 2]
    # v_n = v . n = -0.5 with n = -z, F_damp = 0.2 * 2 m sqrt(kc/m) * v_n, and
    # the z component of -(F_normal + F_damp) n is F_normal + F_damp
    assert dz == pytest.approx(0.2 * 2 * m * math.sqrt(kc / m) * -0.5, rel=1e-10)


def test_friction_stick_and_slip():
    dt = 1e-7
    # stick: tangential trial force below mu_s |Fn|
    d, k = one_node_case(dt, v=(1e-3, 0.0, 0.0), mu=(0.3, 0.2))
    m, kgeo, kmass = stiffness(d, k, dt)
    kc = min(kgeo, kmass)
    cf = d.field("contforce").reshape(-1, 3)[k]
    assert cf[0] == pytest.approx(-kc * 1e-3 * dt, rel=1e-12)
    np.testing.assert_allclose(d.field("ut_prev").reshape(-1, 3)[k], [1e-3 * dt, 0, 0], rtol=1e-14)
    # slip: a large tangential velocity caps at mu_d |Fn| against the motion
    d, k = one_node_case(dt, v=(0.0, -400.0, 0.0), mu=(0.3, 0.2))
    cf = d.field("contforce").reshape(-1, 3)[k]
    assert cf[1] == pytest.approx(0.2 * abs(cf[2]), rel=1e-12)
    assert np.all(d.field("ut_prev").reshape(-1, 3)[k] == 0.0)


def test_outside_triangle_and_first_triangle_rule():
    # a plane whose footprint misses the corner node: no contact
    n = np.array([[0.2, 0.2, 0.0999], [0.5, 0.2, 0.0999], [0.2, 0.5, 0.0999]])
    d, k = one_node_case(1e-7, tri=(n, np.array([[0, 2, 1]])))
    assert np.all(d.field("contforce") == 0.0) and d.field("mesh_in_contact")[k] == -1
    # two planes at different depths: the first listed decides
    n1, t1 = plane(0.1 - 1e-4, flip=True)
    n2, t2 = plane(0.1 - 3e-4, flip=True)
    nodes = np.vstack([n2, n1])
    tris = np.vstack([t2, t1 + 4])
    d1, k = one_node_case(1e-7, delta=3e-4, tri=(nodes, tris))
    d3, _ = one_node_case(1e-7, delta=3e-4)
    assert d1.field("mesh_in_contact")[k] == 0
    np.testing.assert_array_equal(d1.field("contforce"), d3.field("contforce"))


def test_force_persists_after_separation():
    d, k = one_node_case(1e-7)
    before = d.field("contforce").copy()
    x = d.field("x").reshape(-1, 3)
    x[k, 2] = 0.0  # well clear of the plane
    d.setField("x", x.reshape(-1))
    d.CalcContactForces()
    np.testing.assert_array_equal(d.field("contforce"), before)
    assert d.field("mesh_in_contact")[k] == -1
    assert d.contactForce(0) == 0.0


def test_contact_heat_conduction():
    dt, h, Tdie, cp = 1e-7, 5e4, 900.0, 460.0
    d = body(1, 1, 1, 0.1, dt=dt)
    d.setThermal(True, 0.0, cp, 0.9)  # no conduction inside the body, no plastic heat yet
    x = d.field("x").reshape(-1, 3)
    top = np.where(np.isclose(x[:, 2], 0.1))[0]
    k = top[np.argmax(x[top, 0] + x[top, 1])]
    n, t = plane(0.1 - 1e-4, flip=True)
    d.setRigidMesh(n, t)
    d.setContactHeat(h, Tdie)
    d.init()
    xx = x.copy()
    xx[[i for i in top if i != k], 2] = 0.1 - 2e-4
    d.setField("x", xx.reshape(-1))
    T0 = np.full(d.node_count, 20.0)
    T0[k] = 35.0
    d.setField("T", T0)
    d.CalcContactForces()
    q = d.field("q_cont_conv")
    area = d.field("node_area")[k]
    assert area > 0
    assert q[k] == pytest.approx(h * area * (Tdie - 35.0), rel=1e-15)
    assert np.all(np.delete(q, k) == 0.0)
    # the nodal update adds q dt / (m cp): the heat q dt enters as m cp dT
    m = d.field("mdiag")
    d.ThermalCalcs()
    dT = d.field("T") - T0
    assert dT[k] == pytest.approx(q[k] * dt / (m[k] * cp), rel=1e-12)
    assert np.all(np.delete(dT, k) == 0.0)
    assert float(np.sum(m * cp * dT)) == pytest.approx(q[k] * dt, rel=1e-12)
    # a die colder than the node cools it; the flux stays after separation
    d.setContactHeat(h, 0.0)
    d.CalcContactForces()
    q2 = d.field("q_cont_conv")[k]
    assert q2 < 0
    xx[k, 2] = 0.0
    d.setField("x", xx.reshape(-1))
    d.CalcContactForces()
    assert d.field("mesh_in_contact")[k] == -1 and d.field("q_cont_conv")[k] == q2
