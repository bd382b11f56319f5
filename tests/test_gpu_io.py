"""Mesh in from an LS-DYNA keyword file, frames out as legacy VTK (GPU)."""
import numpy as np
import pytest

from oracle.wfref import RefDomain
from weldformfem_amd import Domain, WFX_FP32_MIXED, WFX_FP64, workloads

pytestmark = pytest.mark.gpu


def write_k(path, x, el):
    with open(path, "w") as f:
        f.write("*KEYWORD\n*NODE\n")
        for i, p in enumerate(x):
            f.write(f"{i + 1:8d}{p[0]:16.9e}{p[1]:16.9e}{p[2]:16.9e}       0       0\n")
        f.write("*ELEMENT_SOLID\n")
        for e, nd in enumerate(el):
            f.write(f"{e + 1:8d}       1" + "".join(f"{q + 1:8d}" for q in nd) + "\n")
        f.write("*END\n")


def read_vtk(path):
    """sections of a legacy ASCII VTK file as float arrays"""
    toks = open(path).read().split("\n")
    out, k = {}, 0
    while k < len(toks):
        t = toks[k].split()
        if t and t[0] in ("POINTS", "VECTORS", "SCALARS", "TENSORS", "CELLS", "CELL_TYPES"):
            name = t[0] if t[0] in ("POINTS", "CELLS", "CELL_TYPES") else t[1]
            k += 1
            if t[0] == "SCALARS":
                k += 1  # LOOKUP_TABLE
            vals = []
            while k < len(toks) and toks[k] and (toks[k][0].isdigit() or toks[k][0] in "-+."):
                vals.extend(float(v) for v in toks[k].split())
                k += 1
            out[name] = np.array(vals)
        else:
            k += 1
    return out


def test_keyword_mesh_run_and_vtk_frame(tmp_path):
    print("case: 4x3x2 hexa block through a .k file, 30 steps, VTK frame read back", flush=True)
    ref = RefDomain()
    ref.AddBoxLength([0, 0, 0], [1.0, 0.75, 0.5], 0.125)  # 4 x 3 x 2 cubes of edge 0.25
    x, el = ref.field("x").reshape(-1, 3), ref.field("elnod").reshape(-1, 8)
    kpath = tmp_path / "block.k"
    write_k(kpath, x, el)
    g = Domain(WFX_FP64)
    g.readLSDyna(kpath)
    np.testing.assert_array_equal(g.field("elnod"), el.reshape(-1))
    np.testing.assert_allclose(g.field("x"), x.reshape(-1), rtol=1e-9, atol=1e-12)
    bottom, top = workloads.box_layers(4, 3, 2)
    assert g.sizes() == (5 * 4 * 3, 24, 8)
    workloads.setup_compression(g, bottom, top, -5.0, hg_coeff=0.06)
    g.init()
    g.runFused(30)
    vpath = tmp_path / "frame.vtk"
    g.writeVTK(vpath)
    v = read_vtk(vpath)
    N, E = g.node_count, g.elem_count
    np.testing.assert_allclose(v["POINTS"], g.field("x"), rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(v["DISP"], g.field("u"), rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(v["Velocity"], g.field("v"), rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(v["pl_strain"], g.field("pl_strain"), rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(v["pressure"], g.field("p"), rtol=1e-8, atol=1e-3)
    assert v["CELL_TYPES"].size == E and np.all(v["CELL_TYPES"] == 12)
    cells = v["CELLS"].reshape(E, 9)
    np.testing.assert_array_equal(cells[:, 1:].astype(int), el)
    sig = v["SIGMA"].reshape(E, 3, 3)
    s6 = g.field("sigma").reshape(E, 6)
    np.testing.assert_allclose(sig[:, 0, 1], s6[:, 3], rtol=1e-8, atol=1e-3)
    np.testing.assert_allclose(sig[:, 1, 0], s6[:, 3], rtol=1e-8, atol=1e-3)
    assert v["DISP"].size == 3 * N


def _block(precision):
    ref = RefDomain()
    ref.AddBoxLength([0, 0, 0], [1.0, 0.75, 0.5], 0.125)
    g = Domain(precision)
    g.CreateFromMesh(ref.field("x"), ref.field("elnod"), 8)
    bottom, top = workloads.box_layers(4, 3, 2)
    workloads.setup_compression(g, bottom, top, -5.0, hg_coeff=0.06)
    g.init()
    return g


@pytest.mark.parametrize("precision", [WFX_FP64, WFX_FP32_MIXED])
def test_async_frames_match_sync_frames(tmp_path, precision):
    print(f"case: three asynchronous VTK frames between fused runs, precision {precision}", flush=True)
    g, h = _block(precision), _block(precision)
    for k in range(3):  # the third frame reuses the first slot
        g.runFused(20)
        g.writeVTKAsync(tmp_path / f"a{k}.vtk")
    g.waitVTK()
    for k in range(3):
        h.runFused(20)
        h.writeVTK(tmp_path / f"s{k}.vtk")
    for k in range(3):
        assert (tmp_path / f"a{k}.vtk").read_bytes() == (tmp_path / f"s{k}.vtk").read_bytes(), f"frame {k}"


def test_async_frame_write_error_is_reported(tmp_path):
    from weldformfem_amd._lib import WfxError

    print("case: asynchronous frame into a missing directory", flush=True)
    g = _block(WFX_FP64)
    g.writeVTKAsync(tmp_path / "missing" / "f.vtk")
    with pytest.raises(WfxError, match="cannot write"):
        g.waitVTK()
    g.waitVTK()  # reported once


def test_mesh_change_waits_for_frames_in_flight(tmp_path):
    # the frame writer threads read the host connectivity: a new mesh must
    # wait for them, and the frames must come out whole
    print("case: new mesh right after two asynchronous frames", flush=True)
    g, h = _block(WFX_FP64), _block(WFX_FP64)
    g.runFused(10)
    g.writeVTKAsync(tmp_path / "a0.vtk")
    g.runFused(10)
    g.writeVTKAsync(tmp_path / "a1.vtk")
    g.AddBoxLength([0, 0, 0], [0.1, 0.1, 0.1], 0.025)  # replaces the mesh
    assert g.elem_count == 8
    g.waitVTK()
    for k in range(2):
        h.runFused(10)
        h.writeVTK(tmp_path / f"s{k}.vtk")
        assert (tmp_path / f"a{k}.vtk").read_bytes() == (tmp_path / f"s{k}.vtk").read_bytes(), f"frame {k}"


def expected_vtk(g):
    """the legacy VTK text of g's state, formatted in Python (one thread)"""
    N, E, n = g.node_count, g.elem_count, g.nodxelem
    e9 = lambda v: "%.9e" % v
    vec3 = lambda q: "".join(f"{e9(a)} {e9(b)} {e9(c)}\n" for a, b, c in q.reshape(-1, 3))
    out = ["# vtk DataFile Version 3.0\nwfx explicit step\nASCII\nDATASET UNSTRUCTURED_GRID\n",
           f"POINTS {N} double\n", vec3(g.field("x")), f"CELLS {E} {E * (n + 1)}\n"]
    out += [f"{n} " + " ".join(str(q) for q in row) + "\n" for row in g.field("elnod").reshape(-1, n)]
    out += [f"CELL_TYPES {E}\n", ("12\n" if n == 8 else "10\n") * E, f"POINT_DATA {N}\n"]
    for name, f in [("DISP", "u"), ("Velocity", "v"), ("Acceleration", "a")]:
        out += [f"VECTORS {name} double\n", vec3(g.field(f))]
    out += ["SCALARS nod_mass double 1\nLOOKUP_TABLE default\n", "".join(e9(v) + "\n" for v in g.field("mdiag"))]
    out += [f"CELL_DATA {E}\n"]
    for name, f in [("pressure", "p"), ("pl_strain", "pl_strain"), ("sigma_y", "sigma_y"), ("Vol", "vol")]:
        out += [f"SCALARS {name} double 1\nLOOKUP_TABLE default\n", "".join(e9(v) + "\n" for v in g.field(f))]
    out += ["TENSORS SIGMA double\n"]
    for s in g.field("sigma").reshape(-1, 6):  # xx yy zz xy yz xz
        out += [vec3(np.array([s[0], s[3], s[5], s[3], s[1], s[4], s[5], s[4], s[2]]))]
    return "".join(out).encode()


@pytest.mark.parametrize("sync", [True, False])
def test_frame_text_of_a_larger_block(tmp_path, sync):
    # above the writer's one-thread size: the sections are formatted by several
    # threads and must come out as one thread writes them
    print(f"case: 20^3 block frame (sync={sync}) against a Python rendering of the fields", flush=True)
    ref = RefDomain()
    ref.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], 0.025)
    g = Domain(WFX_FP64)
    g.CreateFromMesh(ref.field("x"), ref.field("elnod"), 8)
    bottom, top = workloads.box_layers(20, 20, 20)
    workloads.setup_compression(g, bottom, top, -5.0, hg_coeff=0.06)
    g.init()
    g.runFused(15)
    p = tmp_path / "f.vtk"
    if sync:
        g.writeVTK(p)
    else:
        g.writeVTKAsync(p)
        g.waitVTK()
    assert p.read_bytes() == expected_vtk(g)
