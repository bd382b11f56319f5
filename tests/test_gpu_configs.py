"""The benchmark workloads at their own sizes, GPU against the oracle.

* 20^3 hexa cube (8,000 elements), top face at -1 m/s, 1000 steps: FP64
  fused and per-step paths within 1e-10 of the oracle at step 100
  (displacement, reaction history, equivalent plastic strain) and within
  1e-8 at step 1000 (rounding differences grow slowly once the block yields);
  FP32-mixed within 1e-4 of the oracle's displacement.
* 64^3-cell tet cylinder (1,239,552 tets), its 500 steps: mesh, connectivity
  and CSR bit for bit; FP64 fused within 1e-10 of the oracle (displacement,
  reaction history, equivalent plastic strain), FP32-mixed within 1e-4 in
  displacement.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def cube(dom):
    dom.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], 0.025)
    bottom, top = workloads.box_layers(20, 20, 20)
    return workloads.setup_compression(dom, bottom, top, -1.0, hg_coeff=0.06)


@pytest.fixture(scope="module")
def cube_oracle():
    r = RefDomain()
    dt = cube(r)
    r.init()
    h100 = r.run(100)
    s100 = {f: r.field(f) for f in ("u", "pl_strain")}
    h = r.run(900)
    s1000 = {f: r.field(f) for f in ("u", "pl_strain")}
    return dt, np.concatenate([h100, h]), s100, s1000


@pytest.mark.parametrize("path", ["fused", "api"])
def test_cube20_1000_steps_fp64(cube_oracle, path):
    dt, hist, s100, s1000 = cube_oracle
    say(f"case: 20^3 hexa cube, 1000 steps, FP64 {path}")
    g = Domain(WFX_FP64)
    assert cube(g) == dt
    g.init()
    run = g.runFused if path == "fused" else g.run
    h100 = run(100)
    e100 = {"u": rel(g.field("u"), s100["u"]), "reaction": rel(h100, hist[:100]),
            "pl": rel(g.field("pl_strain"), s100["pl_strain"])}
    h = np.concatenate([h100, run(900)])
    assert np.max(s1000["pl_strain"]) > 0
    g_u = g.field("u")
    e1000 = {"u": rel(g_u, s1000["u"]), "reaction": rel(h, hist), "pl": rel(g.field("pl_strain"), s1000["pl_strain"])}
    say(f"  rel err at 100 steps {e100}; at 1000 steps {e1000}")
    assert all(v <= 1e-10 for v in e100.values())
    assert all(v <= 1e-8 for v in e1000.values())


def test_cube20_fp32_mixed(cube_oracle):
    dt, hist, s100, s1000 = cube_oracle
    say("case: 20^3 hexa cube, 1000 steps, FP32-mixed vs the FP64 oracle")
    g = Domain(WFX_FP32_MIXED)
    cube(g)
    g.init()
    h = g.runFused(1000)
    eu, er = rel(g.field("u"), s1000["u"]), rel(h, hist)
    ep = rel(g.field("pl_strain"), s1000["pl_strain"])
    say(f"  u {eu:.2e} reaction {er:.2e} pl_strain {ep:.2e}")
    assert eu <= 1e-4 and er <= 1e-3 and ep <= 1e-3


CYL_STEPS = 500


@pytest.fixture(scope="module")
def cylinder_oracle():
    say(f"case: 64^3-cell tet cylinder, oracle {CYL_STEPS} steps")
    x, el, bottom, top = workloads.cylinder_tets(cells=64)
    r = RefDomain()
    r.CreateFromMesh(x, el, 4)
    dt = workloads.setup_compression(r, bottom, top, -1.0, hg_coeff=0.0)
    r.init()
    hist = r.run(CYL_STEPS)
    out = {f: r.field(f) for f in ("u", "pl_strain", "nodel", "nodel_loc", "nodel_offset", "nodel_count", "elnod")}
    return (x, el, bottom, top), dt, hist, out


def test_cylinder_tets_full_size_fp64(cylinder_oracle):
    (x, el, bottom, top), dt, hr, ref = cylinder_oracle
    say(f"case: 64^3-cell tet cylinder, mesh/CSR bit for bit and {CYL_STEPS} FP64 fused steps")
    assert el.shape[0] == 1239552 and x.shape[0] == 218205
    g = Domain(WFX_FP64)
    g.CreateFromMesh(x, el, 4)
    for f in ("nodel", "nodel_loc", "nodel_offset", "nodel_count", "elnod"):
        np.testing.assert_array_equal(g.field(f), ref[f], err_msg=f)
    assert workloads.setup_compression(g, bottom, top, -1.0, hg_coeff=0.0) == dt
    g.init()
    hg = np.concatenate([g.runFused(20), g.runFused(CYL_STEPS - 20)])
    say(f"  largest equivalent plastic strain (oracle) {np.max(ref['pl_strain']):.3e}")
    errs = {"u": rel(g.field("u"), ref["u"]), "reaction": rel(hg, hr),
            "pl_strain": rel(g.field("pl_strain"), ref["pl_strain"])}
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert all(v <= 1e-10 for v in errs.values()), errs


def test_cylinder_tets_full_size_fp32_mixed(cylinder_oracle):
    (x, el, bottom, top), dt, hr, ref = cylinder_oracle
    say(f"case: 64^3-cell tet cylinder, {CYL_STEPS} FP32-mixed fused steps vs the FP64 oracle")
    g = Domain(WFX_FP32_MIXED)
    g.CreateFromMesh(x, el, 4)
    workloads.setup_compression(g, bottom, top, -1.0, hg_coeff=0.0, dt=dt)
    g.init()
    hg = g.runFused(CYL_STEPS)
    eu, er = rel(g.field("u"), ref["u"]), rel(hg, hr)
    ep = rel(g.field("pl_strain"), ref["pl_strain"])
    say(f"  u {eu:.2e} reaction {er:.2e} pl_strain {ep:.2e}")
    assert eu <= 1e-4 and er <= 1e-3 and ep <= 1e-3
