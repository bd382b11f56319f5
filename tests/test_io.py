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
