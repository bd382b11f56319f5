"""The allocation-free energies entry point against wfx_energies, through the
C-ABI.

wfx_energies_noalloc keeps its reduction partials in the library's static
device scratch array instead of a buffer it allocates; it runs the same
kernels, so both entry points return the same bits.  Two calls in a row on
different domains, the second smaller, check that nothing left in the shared
array from the first call reaches the second.  The oracle's computeEnergies
(Mechanical.C:2145-2185) anchors the values.
"""
import ctypes

import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads
from weldformfem_amd._lib import check

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _both(dom):
    """(wfx_energies, wfx_energies_noalloc) of one domain via the C-ABI"""
    out = []
    for fn in (dom._L.wfx_energies, dom._L.wfx_energies_noalloc):
        ek, de = ctypes.c_double(), ctypes.c_double()
        check(fn(dom._h, ctypes.byref(ek), ctypes.byref(de), dom.stream))
        out.append((ek.value, de.value))
    return out


def _domain(cells, precision, steps):
    h = 1.0 / cells
    bottom, top = workloads.box_layers(cells, cells, cells)
    g = Domain(precision)
    g.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], h / 2)
    dt = workloads.setup_compression(g, bottom, top, -1.0)
    g.init()
    g.run(steps)
    r = RefDomain()
    r.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], h / 2)
    workloads.setup_compression(r, bottom, top, -1.0, dt=dt)
    r.init()
    r.run(steps)
    return g, r


@pytest.mark.parametrize("precision", [WFX_FP64, WFX_FP32_MIXED])
def test_energies_noalloc_equals_energies(precision):
    # the first domain fills more of the scratch than the second reads
    for cells, steps in ((24, 30), (6, 20)):
        say(f"case: wfx_energies_noalloc vs wfx_energies, {cells}^3 hexa, {steps} steps, precision {precision}")
        g, r = _domain(cells, precision, steps)
        (ek, de), (ek_n, de_n) = _both(g)
        assert (ek_n, de_n) == (ek, de)  # same kernels, same partial layout: same bits
        ek_r, de_r = r.energies()
        assert ek_r > 0 and de_r != 0
        tol = 1e-12 if precision == WFX_FP64 else 1e-3
        assert ek_n == pytest.approx(ek_r, rel=tol)
        assert de_n == pytest.approx(de_r, rel=tol)
        # a repeated call on the same state returns the same bits again
        assert _both(g)[1] == (ek_n, de_n)
