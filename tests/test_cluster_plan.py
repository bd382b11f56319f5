"""CPU checks of the element-cluster plan the fused element pass uses
(weldformfem_amd/csrc/wfx_cluster.cpp), built host-only:

* the stored order is a permutation of the elements, cut into clusters of at
  most 256 that never span two connected components;
* every node's slots, taken in order and each slot's entries in order, list
  exactly the node's CSR entries (setNodElem, Domain_d.C:1508-1611): each
  (element, local node) once, ascending original element within a slot,
  slots in ascending cluster order;
* each slot names its node and each element's nodes name their slots;
  no cluster has more slots than the element pass stages (a thin column,
  whose 256-element runs touch 1028 nodes, is cut into smaller clusters);
* a structured hexa block is cut into 16 x 4 x 4 bricks;
* a specimen packed with another gets the same clusters and slots (shifted)
  as on its own.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "native", "cluster_plan_check.cpp"),
       os.path.join(ROOT, "weldformfem_amd", "csrc", "wfx_cluster.cpp")]


@pytest.fixture(scope="module")
def planner(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("plan") / "cluster_plan_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread"] + SRC + ["-o", out], check=True, capture_output=True)
    return out


def plan(planner, tmp_path, x, el, n, cs=256):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    el = np.ascontiguousarray(el, dtype=np.int32).reshape(-1, n)
    el.tofile(tmp_path / "el.i32")
    x.tofile(tmp_path / "x.f64")
    pre = str(tmp_path / "p_")
    r = subprocess.run([planner, str(len(x)), str(len(el)), str(n), str(cs), str(tmp_path / "el.i32"),
                        str(tmp_path / "x.f64"), pre], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = {k: np.fromfile(pre + k, dtype=np.int32) for k in
           ("s2o", "eoff", "sbase", "loff", "ent", "nsl_off", "nsl", "snode", "lslot", "spos")}
    _, _, out["max_slots"], out["slot_cap"] = (int(q) for q in r.stdout.split())
    return out


def check_plan(p, el, N, n, cs=256):
    E = len(el)
    s2o = p["s2o"]
    assert np.array_equal(np.sort(s2o), np.arange(E))
    eoff, sbase = p["eoff"], p["sbase"]
    sizes = np.diff(eoff)
    assert eoff[0] == 0 and eoff[-1] == E and np.all(sizes >= 1) and np.all(sizes <= cs)
    C = len(eoff) - 1
    # slot -> (node, [original elements and local nodes])
    slot_node = np.empty(sbase[-1], dtype=np.int64)
    entries = [None] * sbase[-1]
    for c in range(C):
        e0, ne = eoff[c], sizes[c]
        els = s2o[e0:e0 + ne]
        assert np.all(np.diff(els) > 0)  # ascending original order within a cluster
        M = sbase[c + 1] - sbase[c]
        lo = p["loff"][sbase[c] + c: sbase[c] + c + M + 1]
        en = p["ent"][e0 * n:(e0 + ne) * n]
        assert lo[0] == 0 and lo[-1] == ne * n and np.all(np.diff(lo) >= 1)
        for j in range(M):
            q = en[lo[j]:lo[j + 1]]
            t, a = q >> 3, q & 7
            nodes = el[els[t], a]
            assert np.all(nodes == nodes[0])
            assert np.all(np.diff(t) >= 0)
            slot_node[sbase[c] + j] = nodes[0]
            entries[sbase[c] + j] = list(zip(els[t].tolist(), a.tolist()))
    # slot -> node, element-node -> slot of its cluster, and the slot cap
    np.testing.assert_array_equal(p["snode"], slot_node)
    M = np.diff(sbase)
    assert p["max_slots"] == M.max() <= p["slot_cap"]
    cl_of = np.repeat(np.arange(C), sizes)
    ls = p["lslot"].reshape(E, n)
    assert np.all(ls < M[cl_of][:, None])
    np.testing.assert_array_equal(p["snode"][sbase[cl_of][:, None] + ls], el[s2o])
    # each slot's partial sum is stored at its place in the node's slot list
    np.testing.assert_array_equal(p["nsl"][p["spos"]], np.arange(sbase[-1]))
    # the node's slots in cluster order reproduce its CSR list exactly
    ref = RefDomain()
    ref.CreateFromMesh(np.zeros((N, 3)), el.reshape(-1), n)
    off, cnt = ref.field("nodel_offset"), ref.field("nodel_count")
    nodel, loc = ref.field("nodel"), ref.field("nodel_loc")
    nsl_off, nsl = p["nsl_off"], p["nsl"]
    for i in range(N):
        slots = nsl[nsl_off[i]:nsl_off[i + 1]]
        assert np.all(np.diff(slots) > 0) and np.all(slot_node[slots] == i)
        got = sorted(sum((entries[s] for s in slots), []))
        want = sorted(zip(nodel[off[i]:off[i] + cnt[i]].tolist(), loc[off[i]:off[i] + cnt[i]].tolist()))
        assert got == want, i
    return slot_node


def test_hexa_block_bricks(planner, tmp_path):
    r = RefDomain()
    r.AddBoxLength([0, 0, 0], [1.6, 1.6, 0.8], 0.05)  # 16 x 16 x 8 hexa
    x, el = r.field("x"), r.field("elnod").reshape(-1, 8)
    p = plan(planner, tmp_path, x, el, 8)
    check_plan(p, el, len(x) // 3, 8)
    ijk = np.stack(np.unravel_index(np.arange(len(el)), (8, 16, 16)), axis=1)[:, ::-1]  # (i, j, k)
    for c in range(len(p["eoff"]) - 1):
        b = ijk[p["s2o"][p["eoff"][c]:p["eoff"][c + 1]]]
        assert np.array_equal(b.max(axis=0) - b.min(axis=0), [15, 3, 3])
    assert p["sbase"][-1] == (len(el) // 256) * 17 * 5 * 5


@pytest.mark.parametrize("tet", [False, True])
def test_jittered_mesh(planner, tmp_path, tet):
    r = RefDomain()
    h = 0.05
    r.AddBoxLength([0, 0, 0], [7 * h, 5 * h, 6 * h], h / 2, True, tet)
    n = 4 if tet else 8
    x = r.field("x").reshape(-1, 3)
    x = x + (2.0 * workloads.uniform(11, x.size).reshape(-1, 3) - 1.0) * 0.2 * h
    el = r.field("elnod").reshape(-1, n)
    p = plan(planner, tmp_path, x, el, n, cs=64)
    check_plan(p, el, len(x), n, cs=64)


def test_packed_specimen_same_plan(planner, tmp_path):
    r1, r2 = RefDomain(), RefDomain()
    r1.AddBoxLength([0, 0, 0], [0.5, 0.3, 0.4], 0.025)
    r2.AddBoxLength([2.0, 0, 0], [0.3, 0.6, 0.2], 0.025)
    x1, e1 = r1.field("x").reshape(-1, 3), r1.field("elnod").reshape(-1, 8)
    x2, e2 = r2.field("x").reshape(-1, 3), r2.field("elnod").reshape(-1, 8)
    alone = plan(planner, tmp_path, x2, e2, 8)
    x = np.concatenate([x1, x2])
    el = np.concatenate([e1, e2 + len(x1)])
    both = plan(planner, tmp_path, x, el, 8)
    check_plan(both, el, len(x), 8)
    # the second specimen's clusters come after the first's, unchanged
    c1 = int(np.searchsorted(both["eoff"], len(e1)))
    assert both["eoff"][c1] == len(e1)
    np.testing.assert_array_equal(both["eoff"][c1:] - len(e1), alone["eoff"])
    np.testing.assert_array_equal(both["s2o"][len(e1):] - len(e1), alone["s2o"])
    s1 = both["sbase"][c1]
    np.testing.assert_array_equal(both["sbase"][c1:] - s1, alone["sbase"])
    np.testing.assert_array_equal(both["loff"][s1 + c1:], alone["loff"])
    np.testing.assert_array_equal(both["ent"][len(e1) * 8:], alone["ent"])


def test_slot_cap_splits_clusters(planner, tmp_path):
    r = RefDomain()
    r.AddBoxLength([0, 0, 0], [0.05, 0.05, 30.0], 0.025)  # a 1 x 1 x 600 column
    x, el = r.field("x"), r.field("elnod").reshape(-1, 8)
    p = plan(planner, tmp_path, x, el, 8)
    check_plan(p, el, len(x) // 3, 8)
    assert p["slot_cap"] == 512 and p["max_slots"] <= 512
    assert np.diff(p["eoff"]).max() < 256  # 256 elements would touch 1028 nodes
