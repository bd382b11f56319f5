"""Specimen batches, HIP-graph replay and block-size autotuning.

A batch packs independent compression specimens into one mesh with
per-element yield stress and tangent modulus and one reaction group per
specimen.  Each specimen of the batch must follow the oracle run of that
specimen alone with its own material (1e-10 relative after 100 steps in
FP64), and equal, bit for bit, the same specimen run alone on the GPU (the
node gathers and element work never mix specimens).  Graph replays and
autotuned block sizes change only how kernels are launched, so they must
reproduce eager launches bit for bit.
"""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads
from weldformfem_amd._lib import WfxError

pytestmark = pytest.mark.gpu

STATE = ["u", "v", "x", "prev_a", "u_dt", "fi", "pl_strain", "tau", "eps", "p", "sigma_y"]


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def small_batch(count=4):
    # fast, weak specimens: plastic flow within 100 steps
    return workloads.SpecimenBatch(count=count, lo=2, hi=5, seed=3, sy=(100.0e6, 200.0e6), v=(10.0, 30.0))


def batch_domain(b, precision=WFX_FP64, stream=0):
    g = Domain(precision, stream=stream)
    b.setup(g)
    g.init()
    return g


@pytest.mark.parametrize("path", ["fused", "api"])
def test_batch_matches_oracle_per_specimen(path):
    say(f"case: 4-specimen batch, {path} path, vs oracle per specimen")
    b = small_batch()
    g = batch_domain(b)
    assert g.reactionGroupCount() == b.count
    nsteps = 100
    hist = g.runFused(nsteps) if path == "fused" else g.run(nsteps)
    assert hist.shape == (nsteps, b.count)
    u, pl = g.field("u").reshape(-1, 3), g.field("pl_strain")
    plastic = 0
    for s in range(b.count):
        r = RefDomain()
        b.setup_single(r, s)
        r.init()
        rr = r.run(nsteps)
        n0, n1, e0, e1 = b.node_off[s], b.node_off[s + 1], b.elem_off[s], b.elem_off[s + 1]
        eu = rel(u[n0:n1].reshape(-1), r.field("u"))
        er = rel(hist[:, s], rr)
        ep = rel(pl[e0:e1], r.field("pl_strain"))
        say(f"  specimen {s} edges {tuple(b.edges[s])}: u {eu:.2e} reaction {er:.2e} pl {ep:.2e}")
        assert eu <= 1e-10 and er <= 1e-10 and ep <= 1e-10
        plastic += np.max(r.field("pl_strain")) > 0
    assert plastic == b.count
    assert g.nonfiniteCount() == 0


@pytest.mark.parametrize("precision", [WFX_FP64, WFX_FP32_MIXED])
def test_batch_specimen_equals_single_run(precision):
    say(f"case: batch vs single-specimen GPU runs, precision {precision}")
    b = small_batch(3)
    g = batch_domain(b, precision)
    hist = g.runFused(60)
    fields = {f: g.field(f) for f in ["u", "pl_strain", "tau"]}
    for s in range(b.count):
        one = Domain(precision)
        b.setup_single(one, s)
        one.init()
        h1 = one.runFused(60)
        n0, n1, e0, e1 = b.node_off[s], b.node_off[s + 1], b.elem_off[s], b.elem_off[s + 1]
        np.testing.assert_array_equal(hist[:, s], h1)
        np.testing.assert_array_equal(fields["u"].reshape(-1, 3)[n0:n1].reshape(-1), one.field("u"))
        np.testing.assert_array_equal(fields["pl_strain"][e0:e1], one.field("pl_strain"))
        np.testing.assert_array_equal(fields["tau"].reshape(-1, 6)[e0:e1].reshape(-1), one.field("tau"))


@pytest.mark.parametrize("precision", [WFX_FP64, WFX_FP32_MIXED])
def test_graph_replays_match_eager(precision):
    import torch

    say(f"case: graph capture of 15 fused steps, 3 replays vs eager, precision {precision}")
    b = small_batch(3)
    stream = torch.cuda.Stream()
    g = batch_domain(b, precision, stream=stream.cuda_stream)
    e = batch_domain(b, precision)
    g.graphCapture(15)
    for rep in range(3):
        hg = g.graphLaunch()
        torch.cuda.synchronize()
        he = e.runFused(15)
        np.testing.assert_array_equal(hg, he, err_msg=f"replay {rep} reaction history")
        for f in STATE:  # eager reads between the replays
            np.testing.assert_array_equal(g.field(f), e.field(f), err_msg=f"replay {rep} {f}")
    # parameters baked into the graph: a changed time step invalidates it
    g.SetDT(0.5 * g.get("dt"))
    with pytest.raises(WfxError):
        g.graphLaunch()
    g.graphCapture(5)
    g.graphLaunch()


def test_graph_capture_needs_created_stream():
    say("case: graph capture on the null stream is refused")
    g = batch_domain(small_batch(2))
    with pytest.raises(WfxError):
        g.graphCapture(2)


@pytest.mark.parametrize("precision", [WFX_FP64, WFX_FP32_MIXED])
def test_autotune_keeps_state_and_block_sizes_keep_results(precision):
    say(f"case: autotune state preservation and block-size invariance, precision {precision}")
    b = small_batch(3)
    g = batch_domain(b, precision)
    e = batch_domain(b, precision)
    g.runFused(10)
    e.runFused(10)
    before = {f: g.field(f) for f in STATE}
    chosen = g.autotune(reps=3)
    say(f"  chosen block sizes {chosen}")
    assert all(c in (64, 128, 256, 512) for c in chosen)
    for f in STATE:
        np.testing.assert_array_equal(g.field(f), before[f], err_msg=f"autotune changed {f}")
    for sizes in ([64, 512, 128, 64], chosen):
        g.setBlockSizes(sizes)
        np.testing.assert_array_equal(g.runFused(10), e.runFused(10))
        for f in STATE:
            np.testing.assert_array_equal(g.field(f), e.field(f), err_msg=f"{sizes} {f}")
    with pytest.raises(WfxError):
        g.setBlockSizes([256, 256, 96, 256])


def test_reaction_groups_sum_to_single_group():
    say("case: reaction groups vs one group over the same nodes")
    b = small_batch(3)
    g = batch_domain(b)
    g.runFused(20)
    per = g.reactionForce(2)
    g.setReactionNodes(np.concatenate(b.top))
    tot = g.reactionForce(2)
    assert abs(np.sum(per) - tot) <= 1e-10 * np.sum(np.abs(per))
