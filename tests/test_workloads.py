"""Host-side workload generators (no GPU): the packed specimen batch and the
structured hexa box it is made of."""
import numpy as np

from oracle.wfref import RefDomain
from weldformfem_amd import workloads


def test_hexa_box_matches_add_box_length(oracle_lib):
    # AddBoxLength numbering and local order; h = 0.25 keeps the accumulated
    # coordinates exact, so the boxes agree bit for bit
    r = RefDomain()
    r.AddBoxLength([0, 0, 0], [0.75, 0.5, 1.0], 0.125)
    x, el = workloads.hexa_box(3, 2, 4, 0.25)
    np.testing.assert_array_equal(el.reshape(-1), r.field("elnod"))
    np.testing.assert_array_equal(x.reshape(-1), r.field("x"))


def test_specimen_batch_layout():
    b = workloads.SpecimenBatch(count=12, lo=6, hi=40, seed=3)
    assert b.edges.min() >= 6 and b.edges.max() <= 40
    assert np.all((b.sy0 >= 200e6) & (b.sy0 <= 500e6))
    assert np.all((b.Et >= 0.5e9) & (b.Et <= 5e9))
    assert np.all((b.v_top <= -0.5) & (b.v_top >= -3.0))
    ne = np.prod(b.edges, axis=1)
    np.testing.assert_array_equal(np.diff(b.elem_off), ne)
    np.testing.assert_array_equal(np.diff(b.node_off), np.prod(b.edges + 1, axis=1))
    # specimens share no node and their elements only use their own nodes
    for s in range(b.count):
        e0, e1 = b.elem_off[s], b.elem_off[s + 1]
        used = b.elnod[e0:e1]
        assert used.min() == b.node_off[s] and used.max() == b.node_off[s + 1] - 1
        x, el, bottom, top = b.specimen(s)
        assert np.all(x[bottom, 2] == 0.0)
        assert np.allclose(x[top, 2], b.edges[s, 2] * b.h)
    # bounding boxes along x do not overlap
    lo = [b.x[b.node_off[s]:b.node_off[s + 1], 0].min() for s in range(b.count)]
    hi = [b.x[b.node_off[s]:b.node_off[s + 1], 0].max() for s in range(b.count)]
    assert all(hi[s] < lo[s + 1] for s in range(b.count - 1))


def test_specimen_batch_draws_are_reproducible():
    a = workloads.SpecimenBatch(count=5, seed=3)
    b = workloads.SpecimenBatch(count=8, seed=3)
    # the first specimens of a longer batch are the same specimens
    np.testing.assert_array_equal(a.edges, b.edges[:5])
    np.testing.assert_array_equal(a.Et, b.Et[:5])
    c = workloads.SpecimenBatch(count=5, seed=4)
    assert not np.array_equal(a.edges, c.edges)


def test_full_batch_size():
    # the 256-specimen benchmark batch: uneven sizes, about 3M elements
    b = workloads.SpecimenBatch()
    ne = np.diff(b.elem_off)
    assert b.count == 256 and ne.min() < 3000 and ne.max() > 30000
    assert 2.0e6 < b.elem_count < 4.5e6
