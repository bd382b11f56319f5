"""LS-DYNA keyword mesh parser (host side of wfx_read_lsdyna_k), no GPU."""
import os

import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd.domain import parse_lsdyna_k

REF_INPUT = "/root/reference/examples/input"

SMALL_K = """$# two hexa and their nodes, elements before nodes, mixed card styles
*KEYWORD
*ELEMENT_SOLID
$#   eid     pid      n1      n2      n3      n4      n5      n6      n7      n8
       1       1      10      11      14      13      20      21      24      23
2,1,11,12,15,14,21,22,25,24
*NODE
$#   nid               x               y               z      tc      rc
      10             0.0             0.0             0.0       0       0
      11             1.0             0.0             0.0       0       0
      12             2.0             0.0             0.0       0       0
      13             0.0             1.0             0.0       0       0
      14             1.0             1.0             0.0       0       0
      15             2.0             1.0             0.0       0       0
20 0.0 0.0 1.0
21 1.0 0.0 1.0
22 2.0 0.0 1.0
      23             0.0             1.0             1.0       0       0
      24             1.0             1.0             1.0       0       0
      25             2.0             1.0             1.0       0       0
*END
"""

TET_K = """*KEYWORD
*NODE
       7             0.0             0.0             0.0
       5             1.0             0.0             0.0
       9             0.0             1.0             0.0
       8             0.0             0.0             1.0
*ELEMENT_SOLID
       1       1
       7       5       9       8       8       8       8       8
*END
"""


def test_small_hexa_file(tmp_path):
    p = tmp_path / "two.k"
    p.write_text(SMALL_K)
    x, el, n = parse_lsdyna_k(p)
    assert n == 8 and x.shape == (12, 3) and el.shape == (2, 8)
    # ids 10..15, 20..25 in order of appearance -> 0..11
    np.testing.assert_array_equal(el[0], [0, 1, 4, 3, 6, 7, 10, 9])
    np.testing.assert_array_equal(el[1], [1, 2, 5, 4, 7, 8, 11, 10])
    np.testing.assert_array_equal(x[6], [0.0, 0.0, 1.0])
    # unit cubes in the AddBoxLength local order: positive unit volumes
    r = RefDomain()
    r.CreateFromMesh(x, el, 8)
    r.calcElemJAndDerivatives()
    r.CalcElemVol()
    np.testing.assert_allclose(r.field("vol"), [1.0, 1.0], rtol=1e-15)


def test_two_line_tet_card(tmp_path):
    p = tmp_path / "tet.k"
    p.write_text(TET_K)
    x, el, n = parse_lsdyna_k(p)
    assert n == 4
    np.testing.assert_array_equal(el[0], [0, 1, 2, 3])
    np.testing.assert_array_equal(x[1], [1.0, 0.0, 0.0])


def test_errors(tmp_path):
    from weldformfem_amd._lib import WfxError

    p = tmp_path / "bad.k"
    p.write_text("*NODE\n1 0 0 0\n*ELEMENT_SOLID\n1 1 1 2 3 4 5 6 7 8\n")
    with pytest.raises(WfxError, match="unknown node"):
        parse_lsdyna_k(p)
    with pytest.raises(WfxError, match="cannot open"):
        parse_lsdyna_k(tmp_path / "missing.k")


@pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="reference examples not present")
@pytest.mark.parametrize("name,n,E,N", [("cyl_hex.k", 8, 5760, 6479), ("tetra_cyl.k", 4, None, 2513)])
def test_reference_example_meshes(name, n, E, N):
    x, el, nn = parse_lsdyna_k(os.path.join(REF_INPUT, name))
    assert nn == n and x.shape[0] == N
    if E is not None:
        assert el.shape[0] == E
    assert el.min() >= 0 and el.max() < N
    # every node is used and the mesh is a cylinder of radius 12.7 mm or 4.5 mm
    assert np.unique(el).size == N
    rxy = np.hypot(x[:, 0], x[:, 1])
    assert rxy.max() < 0.0128


def test_sparse_ids_crlf_and_lowercase_keywords(tmp_path):
    # ids spread over a wide range take the hash-map path; Windows line ends
    # and lowercase keywords are accepted
    ids = [7, 1_000_003, 42, 99_999_937]  # within the 8-column id field
    lines = ["*keyword", "*node"]
    for k, i in enumerate(ids):
        lines.append(f"{i:8d}{float(k):16.9e}{-0.5 * k:16.9e}{1e-3 * k:16.9e}")
    lines += ["   ", "*element_solid", "       1       1" + "".join(f"{i:8d}" for i in ids) +
              "".join(f"{ids[3]:8d}" * 4), "\t", "*end"]  # blank lines are skipped
    p = tmp_path / "sparse.k"
    p.write_bytes(("\r\n".join(lines) + "\r\n").encode())
    x, el, n = parse_lsdyna_k(p)
    assert n == 4
    np.testing.assert_array_equal(el[0], [0, 1, 2, 3])
    np.testing.assert_array_equal(x[:, 1], [0.0, -0.5, -1.0, -1.5])


def test_duplicate_ids_and_mixed_elements_are_refused(tmp_path):
    from weldformfem_amd._lib import WfxError

    nodes = "*NODE\n" + "".join(f"{i:8d}{0.0:16.9e}{0.0:16.9e}{float(i):16.9e}\n" for i in range(1, 9))
    p = tmp_path / "dup.k"
    p.write_text(nodes + f"{3:8d}{1.0:16.9e}{0.0:16.9e}{0.0:16.9e}\n*ELEMENT_SOLID\n1 1 1 2 3 4 5 6 7 8\n")
    with pytest.raises(WfxError, match="duplicate node id 3"):
        parse_lsdyna_k(p)
    p.write_text(nodes + "*ELEMENT_SOLID\n1 1 1 2 3 4 5 6 7 8\n2 1 1 2 3 4 4 4 4 4\n")
    with pytest.raises(WfxError, match="mixed"):
        parse_lsdyna_k(p)


def test_generated_block_round_trip(tmp_path):
    # a 12x10x8 hexa block written with 9-digit coordinates reads back to the
    # correctly rounded values of the text and the same connectivity
    from weldformfem_amd import workloads

    x, el = workloads.hexa_box(12, 10, 8, 0.0123, origin=(-0.05, 0.1, 0.0))
    p = tmp_path / "block.k"
    with open(p, "w") as f:
        f.write("*KEYWORD\n*NODE\n")
        np.savetxt(f, np.column_stack([np.arange(1, len(x) + 1), x]), fmt="%8d%16.9e%16.9e%16.9e")
        f.write("*ELEMENT_SOLID\n")
        np.savetxt(f, np.column_stack([np.arange(1, len(el) + 1), np.ones(len(el), np.int64), el + 1]),
                   fmt="%8d" * 10)
        f.write("*END\n")
    xs, els, n = parse_lsdyna_k(p)
    assert n == 8
    np.testing.assert_array_equal(els, el)
    want = np.array([float("%.9e" % v) for v in x.reshape(-1)]).reshape(-1, 3)
    np.testing.assert_array_equal(xs, want)
