"""The C++ Domain_d mirror (include/wfx_domain.hpp) driving the library
through the C-ABI gives the same trajectory as the Python binding."""
import os
import subprocess

import numpy as np
import pytest

from weldformfem_amd import Domain, workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_cpp(n, steps, fused):
    exe = os.path.join(ROOT, "weldformfem_amd", "wfx_cube")
    out = subprocess.run([exe, str(n), str(steps), str(int(fused))], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    hist = [float(l.split()[3]) for l in lines if l.startswith("step")]
    energies = [[float(q) for q in l.split()[1:]] for l in lines if l.startswith("energies")]
    last = lines[-1].split()
    run_cpp.energies = energies[0] if energies else None
    return hist, float(last[1]), float(last[3]), float(last[5])


@pytest.mark.parametrize("fused", [False, True])
def test_cpp_driver_matches_python(fused):
    print(f"case: C++ driver 6^3 hexa 200 steps fused={fused}", flush=True)
    hist, dt, umin, epmax = run_cpp(6, 200, fused)
    d = Domain()
    d.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], 0.5 / 6)
    bottom, top = workloads.box_layers(6, 6, 6)
    workloads.setup_compression(d, bottom, top, -1.0, hg_coeff=0.06)
    assert d.get("dt") == pytest.approx(dt, rel=1e-15)
    d.SetDT(dt)
    d.init()
    r = d.runFused(200) if fused else d.run(200)
    np.testing.assert_allclose(hist, r[[99, 199]], rtol=1e-12)
    u = d.field("u").reshape(-1, 3)
    assert u[:, 2].min() == pytest.approx(umin, rel=1e-12)
    assert d.field("pl_strain").max() == pytest.approx(epmax, rel=1e-12, abs=1e-300)
    if not fused:  # computeEnergies through the C++ mirror
        ek, de = d.energies()
        assert run_cpp.energies == pytest.approx([ek, de], rel=1e-12)
