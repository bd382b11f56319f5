"""Synthetic workloads: the compression specimens the benchmark and the tests
run, with a fixed pseudo-random generator so every run is reproducible.

PRNG: SplitMix64.  Draw k (k = 1, 2, ...) of seed s is the SplitMix64 output
for state s + k * 0x9E3779B97F4A7C15; a uniform double is (z >> 11) * 2^-53;
a standard normal is Box-Muller on two consecutive uniforms,
sqrt(-2 ln(1 - u1)) * cos(2 pi u2).

Cube compression: E = 200 GPa, nu = 0.3, rho = 7850 kg/m^3, sigma_y = 300 MPa,
bilinear hardening with tangent modulus 1 GPa, bottom face fixed in z, top
face moving down at a prescribed speed, dt = 0.8 x the critical step (see
cfl_dt).
"""
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)

STEEL = dict(E=200.0e9, nu=0.3, rho=7850.0, sy0=300.0e6, Et=1.0e9)


def splitmix64(seed, n, start=1):
    """n consecutive SplitMix64 outputs (draws start .. start+n-1)."""
    with np.errstate(over="ignore"):
        k = np.arange(start, start + n, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, start=1):
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, n, start=1):
    u = uniform(seed, 2 * n, start)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * math.pi * u2)


# Critical Courant numbers of the Chung-Hulbert scheme (rho_b = 0.8182) on these
# meshes, in units of min edge / dilatational wave speed, measured with the
# oracle (elastic, 2000 steps): 1-point hexa stable at 0.75 and unstable at
# 0.8; six-tet boxes stable at 0.5.  The reference's own rule,
# 0.8 * min edge / sqrt(K/rho) (src/explicit/main.C:876), is 1.02 in these
# units and diverges.
CRITICAL_COURANT = {8: 0.75, 4: 0.4}


def cfl_dt(min_edge, E=STEEL["E"], nu=STEEL["nu"], rho=STEEL["rho"], nodxelem=8, factor=0.8):
    """dt = factor * critical Courant number * min edge / c_d"""
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    cd = math.sqrt((K + 4.0 * G / 3.0) / rho)
    return factor * CRITICAL_COURANT[nodxelem] * min_edge / cd


def box_layers(nx, ny, nz):
    """Node ids of the bottom (z index 0) and top (z index nz) layers of a
    structured (nx+1)(ny+1)(nz+1) node box numbered x fastest."""
    layer = (nx + 1) * (ny + 1)
    bottom = np.arange(layer, dtype=np.int32)
    top = np.arange(nz * layer, (nz + 1) * layer, dtype=np.int32)
    return bottom, top


def cylinder_tets(cells=64, radius=0.5, height=1.0, jitter=0.05, seed=1):
    """Tetrahedral cylinder: a cells^3 grid on [-R,R]^2 x [0,H]; a cell is kept
    when its centre satisfies x^2 + y^2 < R^2; each kept cell is split into six
    tets with the reference's hexa-corner pattern (Domain_d.C:1383-1388).
    Only nodes used by kept cells are numbered (x fastest, then y, then z).
    Every node is moved by U(-jitter*h, +jitter*h) per axis (draws in node
    order, x then y then z).  Returns x[N,3], elnod[E,4], bottom and top
    node ids (grid layers 0 and cells)."""
    n = cells
    hx = 2.0 * radius / n
    hz = height / n
    ci = (np.arange(n) + 0.5) * hx - radius
    cz = np.arange(n)
    X, Y = np.meshgrid(ci, ci, indexing="xy")  # [j, i]
    inside = (X * X + Y * Y) < radius * radius  # [j, i]
    keep = np.broadcast_to(inside, (n, n, n))  # [k, j, i]
    kk, jj, ii = np.nonzero(keep)
    p = n + 1
    gid = lambda i, j, k: (k * p + j) * p + i
    b1 = gid(ii, jj, kk)
    b2 = gid(ii, jj + 1, kk)
    lay = p * p
    corners = np.stack([b1, b1 + 1, b2 + 1, b2, b1 + lay, b1 + lay + 1, b2 + lay + 1, b2 + lay], axis=1)
    pattern = np.array([[0, 1, 3, 5], [1, 2, 3, 5], [0, 5, 3, 4], [4, 5, 3, 7], [5, 6, 3, 7], [5, 2, 3, 6]])
    tets = corners[:, pattern].reshape(-1, 4)
    used = np.zeros(p * p * p, dtype=bool)
    used[tets.reshape(-1)] = True
    old_ids = np.nonzero(used)[0]
    new_id = -np.ones(p * p * p, dtype=np.int64)
    new_id[old_ids] = np.arange(old_ids.size)
    elnod = new_id[tets].astype(np.int32)
    gi = old_ids % p
    gj = (old_ids // p) % p
    gk = old_ids // (p * p)
    x = np.stack([gi * hx - radius, gj * hx - radius, gk * hz], axis=1)
    h = min(hx, hz)
    r = uniform(seed, 3 * old_ids.size).reshape(-1, 3)
    x = x + (2.0 * r - 1.0) * (jitter * h)
    bottom = np.nonzero(gk == 0)[0].astype(np.int32)
    top = np.nonzero(gk == n)[0].astype(np.int32)
    return np.ascontiguousarray(x), np.ascontiguousarray(elnod), bottom, top


def lognormal_sy0(n, median=300.0e6, sigma=0.1, seed=2):
    return median * np.exp(sigma * normal(seed, n))


def hexa_box(nx, ny, nz, h, origin=(0.0, 0.0, 0.0)):
    """Structured hexa box of nx*ny*nz cubes of edge h: nodes x fastest, then
    y, then z, at origin + index*h; local node order of AddBoxLength
    (Domain_d.C:1324-1332): b1, b1+1, b2+1, b2 on the lower layer, then the
    same one layer up.  Returns x[N,3], elnod[E,8]."""
    px, py, pz = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(pz), np.arange(py), np.arange(px), indexing="ij")
    x = np.stack([origin[0] + i.reshape(-1) * h, origin[1] + j.reshape(-1) * h,
                  origin[2] + k.reshape(-1) * h], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    layer = px * py
    b1 = (layer * k + px * j + i).reshape(-1)
    b2 = (layer * k + px * (j + 1) + i).reshape(-1)
    elnod = np.stack([b1, b1 + 1, b2 + 1, b2, b1 + layer, b1 + layer + 1, b2 + layer + 1, b2 + layer], axis=1)
    return np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(elnod, dtype=np.int32)


class SpecimenBatch:
    """Independent hexa compression specimens packed into one mesh.

    Specimen s has edge counts (nx, ny, nz) each drawn uniformly from
    lo..hi, cubes of a common edge h (so every specimen has the same stable
    time step), sigma_y ~ U(sy_lo, sy_hi), tangent modulus ~ U(et_lo, et_hi)
    and top speed ~ U(v_lo, v_hi) (moving down).  Draws are taken six per
    specimen in that order from SplitMix64(seed).  Specimens sit side by side
    along x with a gap and share no node; their nodes and elements are
    contiguous, in specimen order."""

    def __init__(self, count=256, lo=6, hi=40, h=0.025, seed=3, sy=(200.0e6, 500.0e6),
                 et=(0.5e9, 5.0e9), v=(0.5, 3.0)):
        u = uniform(seed, 6 * count).reshape(count, 6)
        self.edges = (lo + np.floor(u[:, :3] * (hi - lo + 1))).astype(np.int64)
        self.sy0 = sy[0] + (sy[1] - sy[0]) * u[:, 3]
        self.Et = et[0] + (et[1] - et[0]) * u[:, 4]
        self.v_top = -(v[0] + (v[1] - v[0]) * u[:, 5])
        self.h = h
        xs, els, self.node_off, self.elem_off = [], [], [0], [0]
        self.bottom, self.top = [], []
        x0 = 0.0
        for s in range(count):
            nx, ny, nz = (int(q) for q in self.edges[s])
            x, el = hexa_box(nx, ny, nz, h, (x0, 0.0, 0.0))
            base = self.node_off[-1]
            b, t = box_layers(nx, ny, nz)
            self.bottom.append(b + base)
            self.top.append(t + base)
            xs.append(x)
            els.append(el + base)
            self.node_off.append(base + x.shape[0])
            self.elem_off.append(self.elem_off[-1] + el.shape[0])
            x0 += (nx + 4) * h
        self.x = np.ascontiguousarray(np.concatenate(xs))
        self.elnod = np.ascontiguousarray(np.concatenate(els))
        self.count = count

    @property
    def elem_count(self):
        return self.elem_off[-1]

    def specimen(self, s):
        """mesh of specimen s alone: x, elnod (local ids), bottom, top"""
        n0, n1 = self.node_off[s], self.node_off[s + 1]
        e0, e1 = self.elem_off[s], self.elem_off[s + 1]
        return (self.x[n0:n1].copy(), self.elnod[e0:e1] - n0, self.bottom[s] - n0, self.top[s] - n0)

    def per_element(self, values):
        return np.repeat(np.asarray(values, dtype=np.float64), np.diff(self.elem_off))

    def dt(self, mat=STEEL):
        return cfl_dt(self.h, mat["E"], mat["nu"], mat["rho"], 8)

    def setup(self, dom, mat=STEEL, hg_coeff=0.06):
        """mesh, per-element sigma_y and tangent modulus, the velocity
        boundary conditions of every specimen and one reaction group per
        specimen (top face) on a Domain"""
        dom.CreateFromMesh(self.x, self.elnod, 8)
        dom.setDensity(mat["rho"])
        dom.setMaterial(0, mat["E"], mat["nu"], mat["rho"], mat["sy0"], mat["Et"], 0.0)
        dom.setElemSy0(self.per_element(self.sy0))
        dom.setElemTangent(self.per_element(self.Et))
        for s in range(self.count):
            for nd in self.bottom[s]:
                dom.AddBCVelNode(int(nd), 2, 0.0)
            for nd in self.top[s]:
                dom.AddBCVelNode(int(nd), 2, float(self.v_top[s]))
        dom.AllocateBCs()
        dom.setReactionGroups(self.top)
        dom.set("hg_coeff", hg_coeff)
        dt = self.dt(mat)
        dom.SetDT(dt)
        return dt

    def setup_single(self, dom, s, mat=STEEL, hg_coeff=0.06):
        """specimen s alone on a Domain or oracle RefDomain, with its own
        material constants: the comparison case of the batch"""
        x, el, bottom, top = self.specimen(s)
        dom.CreateFromMesh(x, el, 8)
        m = dict(mat, sy0=float(self.sy0[s]), Et=float(self.Et[s]))
        return setup_compression(dom, bottom, top, float(self.v_top[s]), m, hg_coeff=hg_coeff, dt=self.dt(mat))


def setup_compression(dom, bottom, top, v_top, mat=STEEL, sy0=None, hg_coeff=0.06, dt=None):
    """Material, boundary conditions, reaction set and time step for a
    compression specimen on a Domain (wfx) or RefDomain (oracle)."""
    dom.setDensity(mat["rho"])
    dom.setMaterial(0, mat["E"], mat["nu"], mat["rho"], mat["sy0"], mat["Et"], 0.0)
    if sy0 is not None:
        dom.setElemSy0(sy0)
    for nd in bottom:
        dom.AddBCVelNode(int(nd), 2, 0.0)
    for nd in top:
        dom.AddBCVelNode(int(nd), 2, float(v_top))
    if hasattr(dom, "AllocateBCs"):
        dom.AllocateBCs()
    dom.setReactionNodes(top)
    dom.set("hg_coeff", hg_coeff)
    if dt is None:
        dt = cfl_dt(dom.calcMinEdgeLength(), mat["E"], mat["nu"], mat["rho"], dom.nodxelem)
    dom.SetDT(dt)
    return dt


# parameter identification: nominal (sigma_y0, Et, hg_coeff) and the hidden
# values the target history is solved at
NOMINAL_PARAMS = (300.0e6, 1.0e9, 0.06)


def hidden_params(seed=4, spread=0.2):
    """nominal parameters scaled by 1 + spread * U(-1, 1), three draws of seed"""
    u = uniform(seed, 3)
    return np.array(NOMINAL_PARAMS) * (1.0 + spread * (2.0 * u - 1.0))
