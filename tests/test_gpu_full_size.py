"""The benchmark workloads at their own sizes.

The 200x200x100 hexa compression block (4,000,000 elements, 4,080,501
nodes, sigma_y0 ~ LogNormal(ln 300 MPa, 0.1)):

* Mesh, connectivity and CSR bit for bit against the oracle, and the first
  100 FP64 fused steps within 1e-10 of it in displacement, equivalent
  plastic strain and reaction history (1e-9 in velocity and pressure, which
  are differences of nearly equal values), the FP32-mixed fused path
  within 1e-4 in displacement (the oracle steps the full block in about
  0.15 s per step on 16 host cores).
* Past yield, where the oracle cannot follow at this size, properties that
  do not depend on it, over the workload's 2000 steps: the fused path and the
  per-step API (two separate kernel sets) agree, and with a uniform yield
  stress the solution is mirror-symmetric about x = 1 and y = 1, within 1e-10
  in displacement and reaction and 1e-9 in plastic strain and pressure;
  FP32-mixed stays within 1e-4 of FP64 in displacement and reaction.

The looser bar on plastic strain and pressure is the conditioning of the
return map, not a different algorithm: dgamma = (sigma_trial - sigma_y) /
(3G + H) subtracts two stresses of ~3e8 Pa, so the 3e-12 relative
differences in strain the two orders of summation leave become ~4e-15 in
dgamma per step, ~4e-12 after a thousand plastic steps: ~4e-10 of the
largest plastic strain (0.9%).  The pressure K (1 - J) amplifies the same
way.

The 256-specimen batch (2,911,675 elements): 1000 FP64 fused steps, four
specimens (the smallest, the largest and two others) equal to the same
specimen run alone, bit for bit; the smallest and largest within 1e-10 of
the oracle (plastic strain 1e-9); FP32-mixed, one captured graph of 1000
steps equal to eager launches, bit for bit.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads

pytestmark = pytest.mark.gpu

NX, NY, NZ = 200, 200, 100
L = [2.0, 2.0, 1.0]


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def block(dom, uniform_sy0=False):
    dom.AddBoxLength([0.0, 0.0, 0.0], L, 0.5 * L[0] / NX)
    bottom, top = workloads.box_layers(NX, NY, NZ)
    sy0 = None if uniform_sy0 else workloads.lognormal_sy0(NX * NY * NZ, seed=2)
    dt = workloads.setup_compression(dom, bottom, top, -2.0, sy0=sy0, hg_coeff=0.06)
    dom.init()
    return dt


ORACLE_STEPS = 100


def test_mesh_and_first_steps_against_the_oracle():
    say(f"case: 200x200x100 block, mesh/CSR bit for bit and {ORACLE_STEPS} FP64 fused steps vs oracle")
    g, r = Domain(WFX_FP64), RefDomain()
    block(g)
    block(r)
    assert (g.node_count, g.elem_count) == (4080501, 4000000)
    for f in ("elnod", "nodel", "nodel_loc", "nodel_offset", "nodel_count"):
        np.testing.assert_array_equal(g.field(f), r.field(f), err_msg=f)
    np.testing.assert_array_equal(g.field("x"), r.field("x"))
    hg = g.runFused(ORACLE_STEPS)
    say("  GPU done, oracle stepping")
    hr = np.concatenate([r.run(ORACLE_STEPS // 2), r.run(ORACLE_STEPS - ORACLE_STEPS // 2)])
    errs = {f: rel(g.field(f), r.field(f)) for f in ("u", "v", "p", "pl_strain")}
    errs["reaction"] = rel(hg, hr)
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # displacement, reaction and plastic strain at 1e-10; velocity and
    # pressure are differences of nearly equal values (p = -K (J_avg - 1)
    # with J_avg - 1 ~ 1e-4), where the hourglass force's other summation
    # order on the fused path shows at ~1e-10: bar 1e-9
    for k, e in errs.items():
        assert e <= (1e-9 if k in ("v", "p") else 1e-10), f"{k}: {e:.3e}"
    del g
    m = Domain(WFX_FP32_MIXED)
    block(m)
    hm = m.runFused(ORACLE_STEPS)
    eu, er = rel(m.field("u"), r.field("u")), rel(hm, hr)
    say(f"  FP32-mixed: u {eu:.2e} reaction {er:.2e}")
    assert eu <= 1e-4 and er <= 1e-3


STEPS = 2000


@pytest.fixture(scope="module")
def fused_run():
    g = Domain(WFX_FP64)
    block(g)
    h = g.runFused(STEPS)
    out = {f: g.field(f) for f in ("u", "pl_strain", "p")}
    return out, h


def test_fused_and_per_step_paths_agree_past_yield(fused_run):
    say(f"case: 200x200x100 block, fused vs per-step API FP64, {STEPS} steps")
    ref, hf = fused_run
    assert np.max(ref["pl_strain"]) > 0  # the block has yielded
    a = Domain(WFX_FP64)
    block(a)
    ha = a.run(STEPS)
    errs = {f: rel(a.field(f), ref[f]) for f in ("u", "pl_strain", "p")}
    errs["reaction"] = rel(ha, hf)
    say("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    for k, e in errs.items():
        assert e <= (1e-9 if k in ("pl_strain", "p") else 1e-10), f"{k}: {e:.3e}"


def test_mirror_symmetry_with_uniform_yield_stress():
    say(f"case: 200x200x100 block, uniform sigma_y0, mirror symmetry after {STEPS} fused steps")
    g = Domain(WFX_FP64)
    block(g, uniform_sy0=True)
    g.runFused(STEPS)
    assert np.max(g.field("pl_strain")) > 0
    u = g.field("u").reshape(NZ + 1, NY + 1, NX + 1, 3)
    scale = np.max(np.abs(u))
    # x -> 2 - x: u_x changes sign, u_y and u_z do not (likewise for y)
    mx = u[:, :, ::-1, :] * np.array([-1.0, 1.0, 1.0])
    my = u[:, ::-1, :, :] * np.array([1.0, -1.0, 1.0])
    ex, ey = np.max(np.abs(u - mx)) / scale, np.max(np.abs(u - my)) / scale
    say(f"  asymmetry x {ex:.2e} y {ey:.2e}")
    assert ex <= 1e-10 and ey <= 1e-10
    ep = g.field("pl_strain").reshape(NZ, NY, NX)
    px = np.max(np.abs(ep - ep[:, :, ::-1])) / np.max(ep)
    py = np.max(np.abs(ep - ep[:, ::-1, :])) / np.max(ep)
    say(f"  plastic strain asymmetry x {px:.2e} y {py:.2e}")
    assert px <= 1e-9 and py <= 1e-9


def test_fp32_mixed_against_fp64(fused_run):
    say(f"case: 200x200x100 block, FP32-mixed vs FP64 after {STEPS} fused steps")
    ref, hf = fused_run
    m = Domain(WFX_FP32_MIXED)
    block(m)
    hm = m.runFused(STEPS)
    eu, er = rel(m.field("u"), ref["u"]), rel(hm, hf)
    say(f"  FP32-mixed rel err: u {eu:.2e} reaction {er:.2e}")
    assert eu <= 1e-4 and er <= 1e-4


# ---- 256 hexa compression specimens (edges U{6..40} per axis, seed 3,
# 2,911,675 elements) packed into one domain, 1000 steps ---------------------
BATCH_STEPS = 1000


def specimen_single(b, s, precision):
    one = Domain(precision)
    b.setup_single(one, s)
    one.init()
    return one


def test_specimen_batch_fp64_against_single_runs_and_oracle():
    say(f"case: 256-specimen batch FP64, {BATCH_STEPS} fused steps, specimens vs single runs and oracle")
    b = workloads.SpecimenBatch()
    g = Domain(WFX_FP64)
    b.setup(g)
    g.init()
    assert g.elem_count == 2911675 and g.reactionGroupCount() == 256
    hist = g.runFused(BATCH_STEPS)
    u, pl = g.field("u").reshape(-1, 3), g.field("pl_strain")
    sizes = np.diff(b.elem_off)
    picks = [int(np.argmin(sizes)), int(np.argmax(sizes)), 100, 200]
    for s in picks:
        n0, n1, e0, e1 = b.node_off[s], b.node_off[s + 1], b.elem_off[s], b.elem_off[s + 1]
        one = specimen_single(b, s, WFX_FP64)
        h1 = one.runFused(BATCH_STEPS)
        np.testing.assert_array_equal(hist[:, s], h1)  # the specimens never mix
        np.testing.assert_array_equal(u[n0:n1].reshape(-1), one.field("u"))
        np.testing.assert_array_equal(pl[e0:e1], one.field("pl_strain"))
        if s in picks[:2]:  # smallest and largest against the oracle
            r = RefDomain()
            b.setup_single(r, s)
            r.init()
            rr = r.run(BATCH_STEPS)
            eu, er = rel(u[n0:n1].reshape(-1), r.field("u")), rel(hist[:, s], rr)
            ep = rel(pl[e0:e1], r.field("pl_strain"))
            say(f"  specimen {s} edges {tuple(b.edges[s])}: u {eu:.2e} reaction {er:.2e} pl {ep:.2e}")
            assert np.max(r.field("pl_strain")) > 0
            assert eu <= 1e-10 and er <= 1e-10 and ep <= 1e-9
    assert g.nonfiniteCount() == 0


def test_specimen_batch_graph_equals_eager_fp32_mixed():
    import torch

    say(f"case: 256-specimen batch FP32-mixed, one graph of {BATCH_STEPS} steps vs eager")
    b = workloads.SpecimenBatch()
    stream = torch.cuda.Stream()
    g = Domain(WFX_FP32_MIXED, stream=stream.cuda_stream)
    e = Domain(WFX_FP32_MIXED)
    for d in (g, e):
        b.setup(d)
        d.init()
    g.graphCapture(BATCH_STEPS)
    hg = g.graphLaunch()
    torch.cuda.synchronize()
    he = e.runFused(BATCH_STEPS)
    np.testing.assert_array_equal(hg, he)
    np.testing.assert_array_equal(g.field("u"), e.field("u"))
    assert np.all(np.isfinite(hg))
