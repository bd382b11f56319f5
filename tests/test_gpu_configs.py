"""The benchmark workloads at their own sizes, GPU against the oracle.

* 20^3 hexa cube (8,000 elements), top face at -1 m/s, 1000 steps: FP64
  fused and per-step paths within 1e-10 of the oracle at step 100 and within
  1e-8 at step 1000 (rounding differences grow slowly once the block yields);
  FP32-mixed within 1e-4 of the oracle's displacement.
* 64^3-cell tet cylinder (1,239,552 tets): mesh, connectivity and CSR bit
  for bit; 20 FP64 steps within 1e-10.
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
    h = np.concatenate([run(100), run(900)])
    assert np.max(s1000["pl_strain"]) > 0
    e100 = max(rel(h[:100], hist[:100]), 0.0)
    g_u = g.field("u")
    e1000 = {"u": rel(g_u, s1000["u"]), "reaction": rel(h, hist), "pl": rel(g.field("pl_strain"), s1000["pl_strain"])}
    say(f"  reaction rel err at 100 steps {e100:.2e}; at 1000 steps {e1000}")
    assert e100 <= 1e-10
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


def test_cylinder_tets_full_size():
    say("case: 64^3-cell tet cylinder, mesh/CSR bit for bit and 20 FP64 steps")
    x, el, bottom, top = workloads.cylinder_tets(cells=64)
    assert el.shape[0] == 1239552 and x.shape[0] == 218205
    g, r = Domain(WFX_FP64), RefDomain()
    for d in (g, r):
        d.CreateFromMesh(x, el, 4)
    for f in ("nodel", "nodel_loc", "nodel_offset", "nodel_count", "elnod"):
        np.testing.assert_array_equal(g.field(f), r.field(f), err_msg=f)
    dt = workloads.setup_compression(g, bottom, top, -1.0, hg_coeff=0.0)
    workloads.setup_compression(r, bottom, top, -1.0, hg_coeff=0.0, dt=dt)
    g.init()
    r.init()
    hg = g.runFused(20)
    hr = r.run(20)
    eu, er = rel(g.field("u"), r.field("u")), rel(hg, hr)
    say(f"  u {eu:.2e} reaction {er:.2e}")
    assert eu <= 1e-10 and er <= 1e-10
