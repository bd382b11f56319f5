"""Stabilisation terms, energies and the error counters on the GPU against the
oracle, through the C-ABI.

* calcArtificialViscosity (Mechanical.C:1948-1977) as a per-step kernel on a
  random state, and inside 100-step trajectories on the per-step and the
  fused paths (the fused element pass adds it after the return map);
* calcElemPressure with every stabilisation term switched on
  (Mechanical.C:691-819), fused path, and its contact branch (alpha_contact,
  hg_coeff_contact, the mean of the viscous terms) in a compression by a
  rigid plane on both paths;
* computeEnergies (Mechanical.C:2145-2185) after per-step steps;
* inverted elements (detJ <= 0, the check of Solver_explicit.C:639-652):
  counted once per element and step like the oracle's count, and reported by
  the run's status.

FP64 bar: 1e-10 relative (1e-14 for one kernel on one state, where the device
pow() may differ from glibc's in the last bit); FP32-mixed 1e-4 on u.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from test_gpu_contact import _pair as contact_pair
from test_gpu_parity import _cyl16, _hexa20, jittered_box, pair, random_state, rel, same
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads
from weldformfem_amd._lib import WfxError

pytestmark = pytest.mark.gpu

AV = dict(artifvisc0=0.6, artifvisc1=0.05)
STAB = dict(alpha_free=0.3, log_factor=0.8, hg_coeff_free=0.2, pspg_scale=0.2, p_pspg_bulkfac=0.05,
            av_coeff_div=0.15, av_coeff_bulk=0.15)


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("tet", [False, True])
def test_artificial_viscosity_kernel(tet):
    say(f"case: calcArtificialViscosity on a random state, tet={tet}")
    x, el, n = jittered_box(5, 4, 3, 0.02, tet, seed=14)
    g, r = pair(x, el, n)
    N, E = g.node_count, g.elem_count
    random_state(g, r, 7, N, E)
    rng = np.random.default_rng(4)
    sr = rng.uniform(-50, 50, (E, 6))
    # traces at, just under and just over the 1e-12 threshold
    sr[0, :3] = [3e-13, 3e-13, 2e-13]
    sr[1, :3] = [4e-13, 4e-13, 3e-13]
    sr[2, :3] = [-5e-13, -5e-13, -5e-13]
    sig = rng.uniform(-3e8, 3e8, 6 * E)
    rho = 7850.0 * rng.uniform(0.9, 1.1, E)
    for d in (g, r):
        d.calcElemJAndDerivatives()
        d.CalcElemVol()
        d.setField("str_rate", sr.reshape(-1))
        d.setField("sigma", sig)
        d.setField("rho", rho)
        for k, v in dict(artifvisc0=40.0, artifvisc1=3.0).items():
            d.set(k, v)
        d.calcArtificialViscosity()
    assert np.any(r.field("sigma") != sig)
    # pow(vol, 1/3) of the device library may differ from glibc's in the last bit
    same("sigma", g.field("sigma"), r.field("sigma"), rtol=1e-14)
    np.testing.assert_array_equal(g.field("sigma").reshape(E, 6)[:1], sig.reshape(E, 6)[:1])


def _run(mesh, nsteps, params, path, precision=WFX_FP64, v_top=-20.0):
    x, el, n, bottom, top = mesh()
    g = Domain(precision)
    g.CreateFromMesh(x, el, n)
    dt = workloads.setup_compression(g, bottom, top, v_top)
    r = RefDomain()
    r.CreateFromMesh(x, el, n)
    workloads.setup_compression(r, bottom, top, v_top, dt=dt)
    for d in (g, r):
        for k, v in params.items():
            d.set(k, v)
        d.init()
    if path == "api":
        hg = g.run(nsteps)
    else:  # several calls: the hand-over between fused runs
        hg = np.concatenate([g.runFused(1), g.runFused(29), g.runFused(nsteps - 30)])
    return g, r, hg, r.run(nsteps)


def _check(g, r, hg, hr, tol=1e-10, fields=("u", "v", "x", "pl_strain", "p", "sigma", "tau")):
    assert np.max(r.field("pl_strain")) > 0
    for f in fields:
        e = rel(g.field(f), r.field(f))
        assert e <= tol, f"{f}: {e:.3e}"
    assert rel(hg, hr) <= tol
    assert g.nonfiniteCount() == 0 and g.invertedCount() == 0


@pytest.mark.parametrize("path", ["api", "fused"])
@pytest.mark.parametrize("mesh", [_hexa20, _cyl16])
def test_artificial_viscosity_trajectory(mesh, path):
    say(f"case: 100 steps with artificial viscosity {AV}, {path} path, {mesh.__name__}")
    g, r, hg, hr = _run(mesh, 100, AV, path)
    _check(g, r, hg, hr)
    # the term is not negligible at these coefficients: a run without it differs
    g0, _, hg0, _ = _run(mesh, 100, {}, path)
    assert rel(hg0, hr) > 1e-8


@pytest.mark.parametrize("mesh", [_hexa20, _cyl16])
def test_artificial_viscosity_fp32_mixed(mesh):
    say(f"case: fused FP32-mixed with artificial viscosity vs FP64 oracle, {mesh.__name__}")
    g, r, hg, hr = _run(mesh, 100, AV, "fused", precision=WFX_FP32_MIXED)
    eu, er = rel(g.field("u"), r.field("u")), rel(hg, hr)
    say(f"  rel err u {eu:.3e} reaction {er:.3e}")
    assert eu <= 1e-4 and er <= 1e-3


@pytest.mark.parametrize("mesh", [_hexa20, _cyl16])
def test_fused_pressure_stabilisation_trajectory(mesh):
    say(f"case: fused 100 steps with every pressure stabilisation term, {mesh.__name__}")
    g, r, hg, hr = _run(mesh, 100, STAB, "fused")
    _check(g, r, hg, hr)


@pytest.mark.parametrize("mesh", [_hexa20, _cyl16])
def test_energies_after_steps(mesh):
    say(f"case: computeEnergies after 40 per-step steps, {mesh.__name__}")
    g, r, _, _ = _run(mesh, 40, AV, "api")
    ek_g, de_g = g.energies()
    ek_r, de_r = r.energies()
    assert ek_r > 0 and de_r != 0
    assert ek_g == pytest.approx(ek_r, rel=1e-12)  # reduction order differs
    assert de_g == pytest.approx(de_r, rel=1e-12)


@pytest.mark.parametrize("path", ["api", "fused"])
def test_contact_pressure_branch(path):
    say(f"case: contact compression with contact-dependent pressure stabilisation, {path} path, 150 steps")
    g, r = contact_pair()
    params = dict(STAB, alpha_contact=0.7, hg_coeff_contact=0.05, **AV)
    for d in (g, r):
        for k, v in params.items():
            d.set(k, v)
    hg = g.run(150) if path == "api" else np.concatenate([g.runFused(13), g.runFused(137)])
    hr = r.run(150)
    assert np.sum(r.field("mesh_in_contact") == 0) > 50
    for f in ["u", "v", "contforce", "pl_strain", "p"]:
        e = rel(g.field(f), r.field(f))
        assert e <= 1e-10, f"{f}: {e:.3e}"
    assert rel(hg, hr) <= 1e-10
    # the branch matters: with the free coefficients in contact too the
    # pressures differ
    g2, r2 = contact_pair()
    for k, v in dict(params, alpha_contact=STAB["alpha_free"], hg_coeff_contact=STAB["hg_coeff_free"]).items():
        g2.set(k, v)
    g2.run(150) if path == "api" else g2.runFused(150)
    assert rel(g2.field("p"), r.field("p")) > 1e-6


@pytest.mark.parametrize("path", ["api", "fused"])
def test_inverted_element_is_reported(path):
    say(f"case: an element pushed inside out, {path} path, FP64 vs oracle")
    r0 = RefDomain()
    r0.AddBoxLength([0, 0, 0], [0.2, 0.2, 0.2], 0.05)
    bottom, top = workloads.box_layers(2, 2, 2)
    x, el = r0.field("x"), r0.field("elnod")
    g, r = Domain(WFX_FP64), RefDomain()
    for d in (g, r):
        d.CreateFromMesh(x, el, 8)
    dt = workloads.setup_compression(r, bottom, top, -1.0, hg_coeff=0.0)
    workloads.setup_compression(g, bottom, top, -1.0, hg_coeff=0.0, dt=dt)
    xs = x.reshape(-1, 3).copy()
    k = int(np.argmax(xs.sum(axis=1)))  # the top corner belongs to one element
    xs[k] -= 0.16
    for d in (g, r):
        d.init()
        d.setField("x", xs.reshape(-1))
    with pytest.raises(WfxError, match="detJ <= 0"):
        g.run(3) if path == "api" else g.runFused(3)
    r.run(3)
    assert r.inverted >= 3  # the element stays inverted: one count per step
    assert g.invertedCount() == r.inverted
    # the reference evaluates h = vol^(1/3) of the inverted element (NaN) in
    # its zero-coefficient pressure and viscosity terms: NaN pressure and
    # stress there, NaN forces scrubbed to zero at its nodes
    assert np.isnan(r.field("p")).any() and np.isnan(r.field("sigma")).any()
    for f in ["u", "v", "p", "sigma"]:
        a, b = g.field(f), r.field(f)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=f)
        ok = ~np.isnan(b)
        e = rel(a[ok], b[ok])
        assert e <= 1e-10, f"{f}: {e:.3e}"
    # check_status 0: the run returns without reading the counters
    g.set("check_status", 0)
    g.run(1) if path == "api" else g.runFused(1)
    assert g.invertedCount() == r.inverted + 1
