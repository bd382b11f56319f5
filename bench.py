"""Benchmark: element-steps/s of the explicit updated-Lagrangian step on one
MI355X, FP64 and FP32-mixed, with the error of the mixed path, the HBM
roofline fraction of the dominant kernel and the oracle timed on the host.

Workload: 2 x 2 x 1 m block of 200 x 200 x 100 reduced-integration hexa
(4,000,000 elements, 4,080,501 nodes), steel, bilinear hardening with a
per-element lognormal yield stress (median 300 MPa, sigma 0.1, seed 2), bottom
face fixed in z, top face at vz = -2 m/s, viscous hourglass 0.06.  A step is
one Chung-Hulbert time step over the whole mesh with all state resident in
HBM.  The headline value is the fused four-kernel step (wfx_step_fused) in
FP64; the per-step Domain_d API path (wfx_run) is timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
I4 = 4


def fused_bytes(n, s, slots_per_elem, clusters_per_elem, hg=True, rho_uniform=True, over8_per_node=0.0,
                f2_lanes=False):
    """Algorithmic bytes per element and per node of the fused kernels: each
    array a kernel touches is moved once, gathered nodal data once per node.
    s = element real size (8 FP64, 4 FP32-mixed); nodal positions are
    double, v, prev_a, u_dt, fi and the ANP ratio are s bytes.  With a uniform density (the
    benchmark's) no density is read.  slots_per_elem = (cluster, node) slots
    per element, clusters_per_elem = clusters per element; over8_per_node =
    share of nodes with more than 8 elements (their CSR offset is read);
    f2_lanes: the nodal-volume pass reads every CSR entry (meshes with more
    than 12 elements per node) instead of the padded first-8 table.
    Both element passes stage their cluster's slots: they read each
    element's nodes as slot indices (uint16 per element-node) and each
    slot's node id (int32), and gather the nodal data once per node.  The
    element pass writes one partial force per slot: FP64 in slot order (the
    node pass reads the node's slot ids), FP32-mixed node-major (it reads
    each slot's position instead, the node pass none)."""
    rho = 1 if (hg and not rho_uniform) else 0
    sn = s
    spe, cpe = slots_per_elem, clusters_per_elem
    # FP64 accumulates the rounding error of the position updates in a float
    # (u from it when a run ends); FP32-mixed recovers u from x alone
    u_rw = 2 * 3 * 4 if s == 8 else 0
    node_major = s == 4
    stage = 2 * n + I4 * spe  # slot indices of the element's nodes, node id per slot
    # tau, eps read and written; plastic strain read and written; sy0;
    # [rho] (pressure and yield stress are stored on a run's last step only)
    f3_reals = 12 + 12 + 2 + 1 + rho
    # staging, element reals, slot entries (uint16 per element-node), slot
    # offsets (uint16 per slot and per cluster), [slot positions], partial
    # forces per slot
    f3_elem = (stage + f3_reals * s + 2 * n + 2 * (spe + cpe) + (I4 * spe if node_major else 0)
               + 3 * s * spe)
    return {
        "f1_volume": (stage + s, 3 * 8),
        # vol gathered (V_n once per run), [rho]; padded first-8 table, count,
        # [offset], V0_n read, mass and ANP ratio written (FP32-mixed stores the
        # ratio's deviation from 1 as a float)
        "f2_nodal": ((s + rho * s + n * I4, 2 * I4 + 8 + 8 + s) if f2_lanes else
                     (s + rho * s, 8 * I4 + I4 + 4 * over8_per_node + 8 + 8 + s)),
        # gathers x, v, ratio
        "f3_element": (f3_elem, 3 * 8 + 3 * sn + s),
        # [slot ids] and partial forces; offsets, mass, BC slot; v, prev_a
        # read and written (the predicted increment is formed from them), x
        # read and written, [the positions' rounding error read and written]
        # (fi, u_dt and u are stored on a run's last step only)
        "f4_node": (spe * ((0 if node_major else I4) + 3 * s),
                    I4 + 8 + I4 + 2 * 2 * 3 * sn + 2 * 3 * 8 + u_rw),
    }


def domain_fused_bytes(d, s, hg=True):
    """fused_bytes for domain d (its slots, clusters and node valences)"""
    N, E, n = d.sizes()
    cnt = d.field("nodel_count")
    return fused_bytes(n, s, d.get("slot_count") / E, d.get("cluster_count") / E, hg=hg,
                       over8_per_node=float(np.mean(cnt > 8)), f2_lanes=n * E > 12 * N)


def api_bytes(n, s, hg=True):
    """the same for the per-step API kernels (one kernel per Domain_d call)"""
    return {
        "update_prediction": (0, 12 * 8),
        "elem_jacobians": (n * I4 + (1 + 3 * n) * s, 3 * 8),
        "elem_vol": (2 * s, 0),
        "nodal_vol": (n * I4 + s, 8 + 2 * I4),
        "nodal_mass": (n * I4 + s, 2 * 8 + 2 * I4),
        "strain_rates": (n * I4 + (1 + 3 * n + 9) * s, 3 * 8),
        "pressure": (2 * n * I4 + 5 * s, 4 * 8 + 2 * I4),
        "stress_strain": (50 * s, 0),
        "elem_forces": ((6 * n + 6) * s, 0),
        "hourglass": ((n * I4 + (2 + 3 * n) * s) if hg else 0, 3 * 8 if hg else 0),
        "assembly": (2 * n * I4 + (3 * n * (2 if hg else 1)) * s, 3 * 8 + 2 * I4),
        "calc_accel": (0, 10 * 8),
        "correction_accvel": (0, 15 * 8),
        "correction_pos": (0, 24 * 8),
    }


def total(bytes_model, E, N):
    return sum(be * E + bn * N for be, bn in bytes_model.values())


def build_domain(precision, nx, ny, nz, L, v_top, stream):
    from weldformfem_amd import Domain, workloads

    d = Domain(precision, stream=stream)
    d.AddBoxLength([0.0, 0.0, 0.0], L, 0.5 * L[0] / nx)
    N, E, n = d.sizes()
    assert (N, E, n) == ((nx + 1) * (ny + 1) * (nz + 1), nx * ny * nz, 8), (N, E, n)
    bottom, top = workloads.box_layers(nx, ny, nz)
    sy0 = workloads.lognormal_sy0(E, seed=2)
    dt = workloads.setup_compression(d, bottom, top, v_top, sy0=sy0, hg_coeff=0.06)
    d.init()
    return d, dt


def timed(fn, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def event_times(calls, reps, torch):
    """average duration (s) of each call, HIP events recorded on the stream
    the kernels are launched on (torch's current stream)"""
    out = {}
    stream = torch.cuda.current_stream()
    for name, fn in calls.items():
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        out[name] = e0.elapsed_time(e1) * 1e-3 / reps
    return out


def copy_bandwidth(torch, gib=2, reps=10):
    """device-to-device copy bandwidth (read + write bytes / time, HIP events):
    torch's copy and the library's streaming copy (16-byte non-temporal
    accesses); returns (best, torch, stream) in bytes/s"""
    from weldformfem_amd import _lib

    L = _lib.load()
    n = (gib << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def rate(fn):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return 2 * n * 8 / (e0.elapsed_time(e1) * 1e-3 / reps)

    def stream_copy():
        _lib.check(L.wfx_stream_copy(src.data_ptr(), dst.data_ptr(), n * 8, stream.cuda_stream))

    r_torch = rate(lambda: dst.copy_(src))
    r_stream = rate(stream_copy)
    assert torch.equal(src, dst)
    del src, dst
    return max(r_torch, r_stream), r_torch, r_stream


PHASES = ("f1_volume", "f2_nodal", "f3_element", "f4_node")


def api_calls(d):
    return {
        "update_prediction": d.UpdatePrediction,
        "elem_jacobians": d.calcElemJAndDerivatives,
        "elem_vol": d.CalcElemVol,
        "nodal_vol": d.CalcNodalVol,
        "nodal_mass": d.CalcNodalMassFromVol,
        "strain_rates": d.calcElemStrainRates,
        "pressure": d.calcElemPressure,
        "stress_strain": lambda: d.CalcStressStrain(0.0),
        "elem_forces": d.calcElemForces,
        "hourglass": d.calcElemHourglassForces,
        "assembly": d.assemblyForces,
        "calc_accel": d.calcAccel,
        "correction_accvel": d.UpdateCorrectionAccVel,
        "correction_pos": d.UpdateCorrectionPos,
    }


def per_kernel_table(times, model, E, N):
    out = {}
    for k, t in times.items():
        be, bn = model[k]
        b = be * E + bn * N
        out[k] = {"us": round(t * 1e6, 2), "GB/s": round(b / t / 1e9, 1), "frac": round(b / t / HBM_PEAK, 3)}
    return out


def cpu_model():
    """the host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds_target=15.0, nx=200, ny=200, nz=100, L=(2.0, 2.0, 1.0)):
    """oracle (reference-faithful C restatement, OpenMP on the host cores) on
    the benchmark block itself: a bounded number of steps right after init,
    in-loop timing"""
    from oracle import wfref
    from weldformfem_amd import workloads

    r = wfref.RefDomain()
    r.AddBoxLength([0.0, 0.0, 0.0], list(L), 0.5 * L[0] / nx)
    E = r.elem_count
    bottom, top = workloads.box_layers(nx, ny, nz)
    workloads.setup_compression(r, bottom, top, -2.0, sy0=workloads.lognormal_sy0(E, seed=2), hg_coeff=0.06)
    r.init()
    t0 = time.perf_counter()
    r.run(2)
    per = (time.perf_counter() - t0) / 2
    n = max(2, min(400, int(seconds_target / max(per, 1e-6))))
    t0 = time.perf_counter()
    r.run(n)
    per = (time.perf_counter() - t0) / n
    return {
        "value": E / per,
        "unit": "element-steps/s",
        "cores": wfref.num_threads(),
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"oracle (C restatement, OpenMP) on the {nx}x{ny}x{nz} hexa block ({E} elements) of the "
                  f"workload, steps 3..{n + 2} after init (elastic), in-loop timing",
    }


def batch_leg(torch, steps, log):
    """256 uneven hexa specimens packed into one FP32-mixed domain: block
    sizes from the autotuner, then `steps` steps as one captured HIP graph"""
    from weldformfem_amd import WFX_FP32_MIXED, Domain, workloads

    b = workloads.SpecimenBatch()
    s = torch.cuda.Stream()  # graph capture needs a created stream
    d = Domain(WFX_FP32_MIXED, stream=s.cuda_stream)
    b.setup(d)
    d.init()
    N, E, _ = d.sizes()
    probe = min(100, steps)
    d.runFused(10, reaction=False)
    t_default = timed(lambda: d.runFused(probe, reaction=False), torch) / probe
    chosen = d.autotune(reps=5)
    t_tuned = timed(lambda: d.runFused(probe, reaction=False), torch) / probe
    log(f"batch: {E} elements, default block sizes {t_default * 1e3:.3f} ms/step, "
        f"autotuned {chosen} {t_tuned * 1e3:.3f} ms/step")
    d.graphCapture(steps, reaction=True)
    d.graphLaunch(reaction=False)  # first replay uploads the graph
    t_graph = timed(lambda: d.graphLaunch(reaction=False), torch) / steps
    hist = d.graphLaunch(reaction=True)
    log(f"batch graph {t_graph * 1e3:.3f} ms/step")
    return {
        "value": E / t_graph,
        "unit": "element-steps/s",
        "workload": f"{b.count} hexa compression specimens, edges U{{6..40}} per axis (seed 3), "
                    f"{E} elements, {N} nodes, one FP32-mixed domain, {steps} steps in one HIP graph",
        "specimen_elements_min_max": [int(min(np.diff(b.elem_off))), int(max(np.diff(b.elem_off)))],
        "steps": steps,
        "ms_per_step_graph": t_graph * 1e3,
        "ms_per_step_eager_default_blocks": t_default * 1e3,
        "ms_per_step_eager_autotuned": t_tuned * 1e3,
        "block_sizes": chosen,
        "reaction_groups": int(hist.shape[1]),
        "reaction_all_finite": bool(np.all(np.isfinite(hist))),
    }


def config_legs(torch, log):
    """The two other single-domain workloads: the 20^3 hexa cube (8,000
    elements, 1000 steps: launch-bound, so eager and HIP-graph launches are
    both timed) and the 64^3-cell tet cylinder (1,239,552 tets, 500 steps)."""
    from weldformfem_amd import WFX_FP32_MIXED, WFX_FP64, Domain, workloads

    out = {}
    s = torch.cuda.Stream()
    for prec, tag in ((WFX_FP64, "fp64"), (WFX_FP32_MIXED, "fp32_mixed")):
        d = Domain(prec, stream=s.cuda_stream)
        d.AddBoxLength([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], 0.025)
        bottom, top = workloads.box_layers(20, 20, 20)
        workloads.setup_compression(d, bottom, top, -1.0, hg_coeff=0.06)
        d.init()
        N, E, _ = d.sizes()
        d.runFused(50, reaction=False)
        t_eager = timed(lambda: d.runFused(1000, reaction=False), torch) / 1000
        d.graphCapture(1000, reaction=False)
        d.graphLaunch(reaction=False)
        t_graph = timed(lambda: d.graphLaunch(reaction=False), torch) / 1000
        out[f"cube20_{tag}"] = {"elements": E, "nodes": N, "us_per_step_eager": t_eager * 1e6,
                                "us_per_step_graph": t_graph * 1e6, "value_graph": E / t_graph}
        log(f"cube 20^3 {tag}: eager {t_eager * 1e6:.1f} us/step, graph {t_graph * 1e6:.1f} us/step")
        del d
    x, el, bottom, top = workloads.cylinder_tets(cells=64)
    for prec, tag in ((WFX_FP64, "fp64"), (WFX_FP32_MIXED, "fp32_mixed")):
        d = Domain(prec, stream=s.cuda_stream)
        d.CreateFromMesh(x, el, 4)
        workloads.setup_compression(d, bottom, top, -1.0, hg_coeff=0.0)
        d.init()
        N, E, _ = d.sizes()
        d.runFused(20, reaction=False)
        bs = d.autotune(reps=5)
        t = timed(lambda: d.runFused(500, reaction=False), torch) / 500
        fb = domain_fused_bytes(d, 8 if prec == WFX_FP64 else 4, hg=False)
        ph = d.timeFusedPhases(50)
        out[f"cylinder_tets_{tag}"] = {"elements": E, "nodes": N, "ms_per_step": t * 1e3, "value": E / t,
                                       "block_sizes": bs,
                                       "step_algorithmic_bytes_per_element": total(fb, E, N) / E,
                                       "step_hbm_frac": total(fb, E, N) / t / HBM_PEAK,
                                       "per_kernel": per_kernel_table(dict(zip(fb, (u * 1e-6 for u in ph))), fb,
                                                                      E, N)}
        log(f"tet cylinder {tag}: {t * 1e3:.3f} ms/step, {E / t:.3e} element-steps/s")
        del d
    return out


# Johnson-Cook (A B n C eps_0 m T_m T_t) and GMT (n1 n2 C1 C2 m1 m2 I1 I2; ranges)
# constants of the extension timings
JC_LAW = (300e6, 500e6, 0.3, 0.014, 1.0, 1.0, 1520.0, 20.0)
GMT_LAW = ((-2e-4, 0.15, 600e6, -1.5e-3, 1e-4, 0.01, -1e-5, 1e-4), (0.002, 3.0, 1e-3, 1e5, 20.0, 1200.0))


def extensions_leg(torch, blocks, steps, log):
    """Step cost of the features around the path, fused, on a fresh copy of
    the benchmark's block (hardening laws, thermal coupling, per-step CFL;
    the benchmark's own domains have been disturbed by the single-kernel
    timings) and on the 64^3-cell tet cylinder compressed by a rigid plane
    with friction (contact, then everything)."""
    from weldformfem_amd import WFX_FP64, Domain, workloads

    out = {}
    for tag, prec, dims, bs in blocks:
        d, _ = build_domain(prec, *dims, -2.0, torch.cuda.current_stream().cuda_stream)
        d.setBlockSizes(bs)
        row = {}
        variants = [("jc", lambda: d.setMaterialJC(200e9, 0.3, 7850.0, *JC_LAW)),
                    ("gmt", lambda: d.setMaterialGMT(200e9, 0.3, 7850.0, 250e6, *GMT_LAW)),
                    ("jc_thermal", lambda: (d.setMaterialJC(200e9, 0.3, 7850.0, *JC_LAW),
                                            d.setThermal(True, 50.0, 460.0, 0.9))),
                    ("jc_thermal_cfl", lambda: d.setCFL(0.3))]
        d.runFused(5, reaction=False)
        row["bilinear"] = timed(lambda: d.runFused(steps, reaction=False), torch) / steps * 1e3
        for name, setup in variants:
            setup()
            d.runFused(5, reaction=False)
            row[name] = timed(lambda: d.runFused(steps, reaction=False), torch) / steps * 1e3
        log(f"extensions {tag} (ms/step): " + ", ".join(f"{k} {v:.3f}" for k, v in row.items()))
        out[f"block_{tag}_ms_per_step"] = row
        del d
    x, el, bottom, top = workloads.cylinder_tets(cells=64)
    zt = float(np.asarray(x).reshape(-1, 3)[:, 2].max())
    die = np.array([[-1, -1, zt], [1, -1, zt], [-1, 1, zt], [1, 1, zt]], float)
    s = torch.cuda.Stream()
    d = Domain(WFX_FP64, stream=s.cuda_stream)
    d.CreateFromMesh(x, el, 4)
    d.setDensity(7850.0)
    d.setMaterial(0, 200e9, 0.3, 7850.0, 300e6, 1e9, 0.0)
    for nd in bottom:
        d.AddBCVelNode(int(nd), 2, 0.0)
    d.AllocateBCs()
    d.SetDT(workloads.cfl_dt(d.calcMinEdgeLength(), 200e9, 0.3, 7850.0, 4))
    d.init()
    row = {}
    d.runFused(5, reaction=False)
    row["plain"] = timed(lambda: d.runFused(steps, reaction=False), torch) / steps * 1e3
    d.setRigidMesh(die, np.array([[0, 2, 1], [1, 2, 3]]), vel=[0.0, 0.0, -2.0])
    d.setContactParams(0.1, 0.2, 0.1)
    d.runFused(5, reaction=False)
    row["contact"] = timed(lambda: d.runFused(steps, reaction=False), torch) / steps * 1e3
    d.setMaterialJC(200e9, 0.3, 7850.0, *JC_LAW)
    d.setThermal(True, 50.0, 460.0, 0.9)
    d.setCFL(0.25)
    d.runFused(5, reaction=False)
    row["contact_jc_thermal_cfl"] = timed(lambda: d.runFused(steps, reaction=False), torch) / steps * 1e3
    N, E, _ = d.sizes()
    log("extensions tet cylinder + die (ms/step): " + ", ".join(f"{k} {v:.3f}" for k, v in row.items()))
    out["cylinder_tets_die_fp64_ms_per_step"] = dict(row, elements=E, nodes=N,
                                                      contact_force_N=d.contactForce(0))
    del d
    return out


def large_leg(torch, log, nx=400, ny=400, nz=200, steps=50):
    """The same compression block eight times larger (400x400x200, 32M
    elements): device memory per element and the fused step rate, FP64 and
    FP32-mixed, autotuned"""
    from weldformfem_amd import WFX_FP32_MIXED, WFX_FP64

    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": f"hexa compression block {nx}x{ny}x{nz} ({nx * ny * nz} elements), fused step"}
    for prec, tag in ((WFX_FP64, "fp64"), (WFX_FP32_MIXED, "fp32_mixed")):
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        t0 = time.perf_counter()
        d, _ = build_domain(prec, nx, ny, nz, [nx / 100.0, ny / 100.0, nz / 100.0], -2.0, stream)
        t_setup = time.perf_counter() - t0
        d.runFused(3, reaction=False)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        bs = d.autotune(reps=5)
        d.runFused(5, reaction=False)
        t = timed(lambda: d.runFused(steps, reaction=False), torch) / steps
        E = nx * ny * nz
        out[tag] = {"value": E / t, "ms_per_step": t * 1e3, "device_bytes_per_element": (free0 - free1) / E,
                    "setup_s": t_setup, "block_sizes": bs}
        log(f"large {tag}: {t * 1e3:.3f} ms/step, {E / t:.3e} element-steps/s, "
            f"{(free0 - free1) / E:.0f} B/element on the device, set-up {t_setup:.1f} s")
        del d
        torch.cuda.synchronize()
    return out


def write_k_file(path, x, elnod):
    """LS-DYNA keyword file (*NODE fixed 8+16+16+16, *ELEMENT_SOLID 8+8 then
    8-wide node ids), ids from 1"""
    N, E = x.shape[0], elnod.shape[0]
    with open(path, "w") as f:
        f.write("*KEYWORD\n*NODE\n")
        np.savetxt(f, np.column_stack([np.arange(1, N + 1), x]), fmt="%8d%16.9e%16.9e%16.9e")
        f.write("*ELEMENT_SOLID\n")
        np.savetxt(f, np.column_stack([np.arange(1, E + 1), np.ones(E, np.int64), elnod + 1]),
                   fmt="%8d" * (2 + elnod.shape[1]))
        f.write("*END\n")


def io_leg(torch, log, nx=100, ny=100, nz=50, steps=200):
    """Mesh in and frames out on a 100x100x50 hexa compression block (500,000
    elements): the keyword-file parser and the whole mesh set-up from a file;
    a synchronous VTK frame; an asynchronous one: how long the caller's stream
    is held (the device snapshot), the step rate while a host thread writes
    the file against the rate without a frame, and the wait for the file."""
    import shutil
    import tempfile

    from weldformfem_amd import WFX_FP64, Domain, workloads
    from weldformfem_amd.domain import parse_lsdyna_k

    tmp = tempfile.mkdtemp(prefix="wfx_io_")
    out = {"workload": f"{nx}x{ny}x{nz} hexa block ({nx * ny * nz} elements), FP64, files in a temporary "
                       "directory of the host"}
    try:
        h = 1.0 / nx
        x, el = workloads.hexa_box(nx, ny, nz, h)
        kpath = os.path.join(tmp, "block.k")
        write_k_file(kpath, x, el)
        mb = os.path.getsize(kpath) / 1e6
        t0 = time.perf_counter()
        xs, els, n = parse_lsdyna_k(kpath)
        t_parse = time.perf_counter() - t0
        assert n == 8 and els.shape[0] == el.shape[0] and np.array_equal(els, el)
        d = Domain(WFX_FP64, stream=torch.cuda.current_stream().cuda_stream)
        t0 = time.perf_counter()
        d.readLSDyna(kpath)
        torch.cuda.synchronize()
        t_read = time.perf_counter() - t0
        out["k_file_MB"] = mb
        out["k_parse_s"] = t_parse
        out["k_parse_MB_per_s"] = mb / t_parse
        out["k_parse_elements_per_s"] = el.shape[0] / t_parse
        out["k_read_and_setup_s"] = t_read
        log(f"io: .k {mb:.1f} MB parsed in {t_parse:.3f} s, read + device set-up {t_read:.3f} s")
        bottom, top = workloads.box_layers(nx, ny, nz)
        workloads.setup_compression(d, bottom, top, -2.0)
        d.init()
        d.runFused(20, reaction=False)
        t_ref = timed(lambda: d.runFused(steps, reaction=False), torch) / steps
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.writeVTK(os.path.join(tmp, "sync.vtk"))
        t_sync = time.perf_counter() - t0
        vtk_mb = os.path.getsize(os.path.join(tmp, "sync.vtk")) / 1e6
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.writeVTKAsync(os.path.join(tmp, "async.vtk"))
        t_call = time.perf_counter() - t0
        torch.cuda.synchronize()
        t_snap = time.perf_counter() - t0
        t0 = time.perf_counter()
        d.runFused(steps, reaction=False)
        torch.cuda.synchronize()
        t_during = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        d.waitVTK()
        t_wait = time.perf_counter() - t0
        out.update({"vtk_file_MB": vtk_mb, "vtk_sync_s": t_sync, "vtk_async_call_s": t_call,
                    "vtk_async_snapshot_s": t_snap, "ms_per_step_no_frame": t_ref * 1e3,
                    "ms_per_step_while_writing": t_during * 1e3, "steps_while_writing": steps,
                    "vtk_async_wait_after_steps_s": t_wait})
        log(f"io: VTK {vtk_mb:.1f} MB sync {t_sync:.3f} s; async call {t_call * 1e3:.2f} ms, snapshot "
            f"{t_snap * 1e3:.2f} ms, steps {t_ref * 1e3:.3f} -> {t_during * 1e3:.3f} ms while writing, "
            f"wait {t_wait:.3f} s")
        del d
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def adjoint_leg(torch, log, steps=2000, cells=30, with_oracle=True):
    """Parameter identification workload: 30^3 hexa compression, reaction
    history loss against a target solved at hidden parameters; adjoint
    gradient (checkpoint every 50 steps) timed and checked against central
    differences of the GPU forward and of the oracle."""
    from weldformfem_amd import WFX_FP32_MIXED, WFX_FP64, workloads
    from weldformfem_amd.adjoint import CompressionProblem, identify

    hidden = workloads.hidden_params(seed=4)        # nominal x (1 + 0.2 U(-1, 1))
    theta0 = np.array(workloads.NOMINAL_PARAMS)     # identification starts at the nominal values
    pb = CompressionProblem(cells=cells, nsteps=steps, every=50)
    target = pb.forward(hidden)
    E = pb.domain.elem_count
    pb.forward(theta0)
    torch.cuda.synchronize()
    t_fwd = timed(lambda: pb.forward(theta0), torch)
    r = pb.forward(theta0) - target
    L = float(np.dot(r, r))
    t_bwd = timed(lambda: pb.backward(2.0 * r), torch)
    g = pb.backward(2.0 * r)

    def gpu_loss(t):
        d = pb.forward(t) - target
        return float(np.dot(d, d))

    # relative difference steps: 1e-6 where the loss is sensitive; hg_coeff
    # moves the loss by ~2.5e-4 of itself and its loss is piecewise smooth
    # (elements cross the yield surface at slightly different steps), so its
    # difference converges to the adjoint only as the step falls to ~3e-5,
    # below which the loss's rounding takes over (hg_fd_sweep shows the trend)
    steps_rel = [1e-6 if abs(t * gi) >= 0.1 * L else 3e-5 for t, gi in zip(theta0, g)]

    def cfd(loss):
        out = np.zeros(3)
        for i in range(3):
            h = steps_rel[i] * theta0[i]
            tp, tm = theta0.copy(), theta0.copy()
            tp[i] += h
            tm[i] -= h
            out[i] = (loss(tp) - loss(tm)) / (2 * h)
        return out

    fd = cfd(gpu_loss)
    log(f"adjoint: grad {g}, GPU fd {fd}")
    sweep = []
    for hr in (3e-4, 1e-4, 3e-5):
        h = hr * theta0[2]
        tp, tm = theta0.copy(), theta0.copy()
        tp[2] += h
        tm[2] -= h
        fdh = (gpu_loss(tp) - gpu_loss(tm)) / (2 * h)
        sweep.append({"rel_step": hr, "fd": fdh, "rel_err": abs(g[2] - fdh) / abs(fdh)})
    out = {
        "workload": f"{cells}^3 hexa compression ({E} elements), {steps} steps, loss = sum (R - R*)^2 over the "
                    "top-face reaction history, theta = (sigma_y0, Et, hg_coeff), checkpoint every 50 steps",
        "forward_record_s": t_fwd, "backward_s": t_bwd,
        "backward_over_forward": t_bwd / t_fwd,
        "grad": g.tolist(), "fd_rel_steps": steps_rel,
        "fd_gpu": fd.tolist(), "rel_err_vs_fd_gpu": (np.abs(g - fd) / np.abs(fd)).tolist(),
        "hg_fd_sweep_gpu": sweep,
    }
    if with_oracle:
        from oracle import wfref

        def oracle_hist(t):
            ref = wfref.RefDomain()
            ref.AddBoxLength([0, 0, 0], [1.0, 1.0, 1.0], 0.5 / cells)
            bottom, top = workloads.box_layers(cells, cells, cells)
            m = dict(workloads.STEEL, sy0=float(t[0]), Et=float(t[1]))
            workloads.setup_compression(ref, bottom, top, -1.0, m, hg_coeff=float(t[2]))
            ref.init()
            return ref.run(steps)

        t0 = time.perf_counter()
        tgt = oracle_hist(hidden)

        def oracle_loss(t):
            d = oracle_hist(t) - tgt
            return float(np.dot(d, d))

        fdo = cfd(oracle_loss)
        out["fd_oracle"] = fdo.tolist()
        out["rel_err_vs_fd_oracle"] = (np.abs(g - fdo) / np.abs(fdo)).tolist()
        out["oracle_seconds"] = time.perf_counter() - t0
        log(f"adjoint: oracle fd {fdo}")
    # FP32-mixed forward with the FP64 adjoint
    q = CompressionProblem(cells=cells, nsteps=steps, every=50, precision=WFX_FP32_MIXED)
    L32, g32 = q.loss_and_grad(theta0, target)
    out["fp32_mixed"] = {"grad": g32.tolist(), "rel_err_vs_fp64": (np.abs(g32 - g) / np.abs(g)).tolist(),
                         "loss_rel_err": abs(L32 - L) / L}
    # identification from theta0
    t0 = time.perf_counter()
    theta, loss, nit, losses = identify(pb, target, theta0, scale=np.array([300e6, 1e9, 0.06]), max_iter=100)
    out["identification"] = {"theta0": theta0.tolist(), "hidden": hidden.tolist(), "found": theta.tolist(),
                             "rel_err": (np.abs(theta - hidden) / hidden).tolist(), "iterations": nit,
                             "loss_ratio": loss / losses[0], "seconds": time.perf_counter() - t0}
    log(f"adjoint: identification {out['identification']}")
    return out


def pmc_probe(args):
    """Kernels for hardware-counter runs (rocprofv3 --pmc): a plain 1 GiB
    device-to-device copy whose FETCH_SIZE/WRITE_SIZE calibrate the counters,
    then, on the benchmark mesh at the given block sizes, the headline's
    window: --pre-steps fused steps (past first yield), then a run of five
    fused steps of each path (the inner three dispatches of each kernel are
    the ones summarised: a run's first and last node passes move extra
    bytes) and one per-step API step."""
    import torch

    from weldformfem_amd import WFX_FP32_MIXED, WFX_FP64

    torch.cuda.init()
    stream = torch.cuda.current_stream().cuda_stream
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    nx, ny, nz, L = (200, 200, 100, [2.0, 2.0, 1.0]) if not args.quick else (100, 100, 50, [2.0, 2.0, 1.0])
    for prec, bs in ((WFX_FP64, args.bs64), (WFX_FP32_MIXED, args.bs32)):
        d, _ = build_domain(prec, nx, ny, nz, L, -2.0, stream)
        if bs:
            d.setBlockSizes([int(q) for q in bs.split(",")])
        d.runFused(args.pre_steps, reaction=False)
        d.runFused(5, reaction=False)
        if prec == WFX_FP64:
            d.run(1, reaction=False)
        torch.cuda.synchronize()
        N, E, n = d.sizes()
        model = domain_fused_bytes(d, 8 if prec == WFX_FP64 else 4)
        print("PMC_MODEL " + json.dumps({"precision": "fp64" if prec == WFX_FP64 else "fp32_mixed",
                                         "bytes": {k: be * E + bn * N for k, (be, bn) in model.items()}}),
              flush=True)
        print(f"pmc probe: precision {prec} done, largest plastic strain {d.field('pl_strain').max():.3e}",
              flush=True)
        del d
    print("pmc probe done", flush=True)


def pmc_summary(fetch_csv, write_csv, probe_log):
    """Markdown table of HBM traffic per dispatch of the fused kernels from two
    rocprofv3 --pmc runs of --pmc-probe (FETCH_SIZE, WRITE_SIZE; KiB per
    dispatch): each counter scaled by what it reads for the probe's 1 GiB
    device copy, the median over the inner three of each kernel's last five
    dispatches (the probe's last run), against the byte model the probe
    printed."""
    import csv
    import statistics

    def load(path, counter):
        by = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] == counter:
                    by.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return by

    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    gib_kib = float(1 << 20)
    cf = gib_kib / max(v for v in fetch["__amd_rocclr_copyBuffer"])  # KiB read for a 1 GiB copy
    cw = gib_kib / max(v for v in write["__amd_rocclr_copyBuffer"])
    models = {}
    with open(probe_log) as f:
        for line in f:
            if line.startswith("PMC_MODEL "):
                m = json.loads(line[len("PMC_MODEL "):])
                models[m["precision"]] = m["bytes"]
    lines = [f"Calibration: the probe's 1 GiB device copy reads {gib_kib / cf:.0f} KiB of FETCH_SIZE and "
             f"{gib_kib / cw:.0f} KiB of WRITE_SIZE: reads = {cf:.3f} x FETCH_SIZE, writes = {cw:.3f} x WRITE_SIZE.",
             "Per-dispatch values: median of the inner three of each kernel's last five dispatches (a five-step "
             "run past first yield; its first and last node passes move extra bytes).", "",
             "| kernel | reads MB | writes MB | HBM MB | model MB | HBM / model |", "|---|---:|---:|---:|---:|---:|"]
    out = {}
    kernels = {"f1_volume": "f1_cluster<", "f2_nodal": "f2_nodal<", "f3_element": "f3_cluster<",
               "f4_node": "f4_node<"}
    for name in fetch:
        short = next((k for k, tag in kernels.items() if tag in name), None)
        if short is None or name not in write:
            continue
        prec = "fp64" if "<double" in name else "fp32_mixed"
        rd = statistics.median(fetch[name][-4:-1]) * cf * 1024 / 1e6
        wr = statistics.median(write[name][-4:-1]) * cw * 1024 / 1e6
        mb = models.get(prec, {}).get(short)
        ratio = (rd + wr) / (mb / 1e6) if mb else float("nan")
        out.setdefault(prec, {})[short] = (rd + wr) * 1e6
        label = name.replace("void wfx::(anonymous namespace)::", "").split(">(")[0] + ">"
        lines.append(f"| `{label}` | {rd:.1f} | "
                     f"{wr:.1f} | {rd + wr:.1f} | {mb / 1e6 if mb else float('nan'):.1f} | {ratio:.2f} |")
    print("\n".join(lines))
    print("\nPMC_TRAFFIC " + json.dumps(out))


def sq_summary(csvs):
    """Markdown table of SQ_* counters of the fused passes from rocprofv3
    --pmc runs of --pmc-probe: per kernel the median over the inner three of
    its last five dispatches; SQ_INSTS_* per wavefront, the cycle counters as
    fractions of SQ_BUSY_CYCLES or SQ_WAVE_CYCLES as named."""
    import csv
    import statistics

    vals = {}
    for path in csvs:
        with open(path) as f:
            for row in csv.DictReader(f):
                name = row["Kernel_Name"]
                if not any(t in name for t in ("f1_cluster<", "f2_nodal<", "f3_cluster<", "f4_node<")):
                    continue
                label = name.replace("void wfx::(anonymous namespace)::", "").split(">(")[0] + ">"
                vals.setdefault(label, {}).setdefault(row["Counter_Name"], []).append(float(row["Counter_Value"]))
    counters = sorted({c for d in vals.values() for c in d})
    lines = ["| kernel | " + " | ".join(counters) + " |", "|---|" + "---:|" * len(counters)]
    for label, d in vals.items():
        med = {c: statistics.median(v[-4:-1]) for c, v in d.items() if len(v) >= 4}
        waves = med.get("SQ_WAVES", 0.0)
        cells = []
        for c in counters:
            v = med.get(c)
            if v is None:
                cells.append("")
            elif c.startswith("SQ_INSTS") and waves > 0:
                cells.append(f"{v / waves:.1f}/wave")
            else:
                cells.append(f"{v:.4g}")
        lines.append(f"| `{label}` | " + " | ".join(cells) + " |")
    print("\n".join(lines))


def stats_summary(stats_csv, trace_csv, rows=24, window=25):
    """Markdown table of a rocprofv3 --kernel-trace --stats run: the `rows`
    kernels with the most total time, and, per kernel, the mean duration of
    its last `window` dispatches from the trace (for the fused passes: the
    plastic window the benchmark times)."""
    import csv

    def label(name):
        return name.replace("void wfx::(anonymous namespace)::", "").replace("wfx::(anonymous namespace)::", "") \
            .split(">(")[0].split("(")[0] + (">" if ">(" in name else "")

    durations = {}
    with open(trace_csv) as f:
        for row in csv.DictReader(f):
            durations.setdefault(row["Kernel_Name"], []).append(
                (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3)
    lines = [f"| kernel | calls | avg (us) | min (us) | max (us) | total (ms) | % | last {window} avg (us) |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]
    with open(stats_csv) as f:
        for k, row in enumerate(csv.DictReader(f)):
            if k >= rows:
                break
            d = durations.get(row["Name"], [])
            tail = sum(d[-window:]) / len(d[-window:]) if d else float("nan")
            lines.append(f"| `{label(row['Name'])}` | {row['Calls']} | {float(row['AverageNs']) / 1e3:.2f} | "
                         f"{float(row['MinNs']) / 1e3:.2f} | {float(row['MaxNs']) / 1e3:.2f} | "
                         f"{float(row['TotalDurationNs']) / 1e6:.2f} | {float(row['Percentage']):.2f} | {tail:.2f} |")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--err-steps", type=int, default=2000, help="steps after which the mixed error is taken")
    ap.add_argument("--quick", action="store_true", help="smaller mesh (100x100x50), no CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-api", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-extensions", action="store_true")
    ap.add_argument("--no-io", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--bs64", default="", help="fused block sizes for FP64 (f1,f2,f3,f4) instead of autotuning")
    ap.add_argument("--bs32", default="", help="the same for FP32-mixed")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--pre-steps", type=int, default=1000,
                    help="fused steps before the headline's warmup and timed steps (the block yields near step 730)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "PROBE_LOG"),
                    help="print the traffic table of two --pmc-probe counter runs")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="short run for counter collection: a 1 GiB device copy (calibration), "
                         "then a 5-step fused run FP64 and FP32-mixed and one per-step API step")
    ap.add_argument("--sq-summary", nargs="+", metavar="PMC_CSV",
                    help="print the per-wavefront SQ counter table of --pmc-probe counter runs")
    ap.add_argument("--stats-summary", nargs=2, metavar=("KERNEL_STATS_CSV", "KERNEL_TRACE_CSV"),
                    help="print the kernel table of a rocprofv3 --kernel-trace --stats run")
    args = ap.parse_args()
    if args.pmc_summary:
        return pmc_summary(*args.pmc_summary)
    if args.stats_summary:
        return stats_summary(*args.stats_summary)
    if args.sq_summary:
        return sq_summary(args.sq_summary)
    if args.pmc_probe:
        return pmc_probe(args)
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is a design only)", file=sys.stderr)
        sys.exit(2)

    import torch

    from weldformfem_amd import WFX_FP32_MIXED, WFX_FP64

    torch.cuda.init()
    stream = torch.cuda.current_stream().cuda_stream
    nx, ny, nz, L = (200, 200, 100, [2.0, 2.0, 1.0]) if not args.quick else (100, 100, 50, [2.0, 2.0, 1.0])
    copy_bw, copy_torch, copy_stream = copy_bandwidth(torch)

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    log(f"building {nx}x{ny}x{nz} hexa FP64")
    d64, dt = build_domain(WFX_FP64, nx, ny, nz, L, -2.0, stream)
    N, E, n = d64.sizes()
    if args.bs64:
        bs64 = [int(q) for q in args.bs64.split(",")]
        d64.setBlockSizes(bs64)
    else:
        bs64 = d64.autotune(reps=5)
    log(f"FP64 block sizes {'(given)' if args.bs64 else '(autotuned)'} {bs64}")
    # the timed window lies past first yield (near step 730 at 2 m/s), where
    # the return map and the plastic-strain stores run
    d64.runFused(args.pre_steps, reaction=False)
    steps64 = args.pre_steps
    d64.runFused(args.warmup, reaction=False)
    steps64 += args.warmup
    yielded64 = float(np.mean(d64.field("pl_strain") > 0))
    s64 = timed(lambda: d64.runFused(args.steps, reaction=False), torch) / args.steps
    steps64 += args.steps
    log(f"FP64 fused {s64 * 1e3:.3f} ms/step")
    # per-pass durations in the run's own order (HIP events between the
    # passes on the stream they run on), in the same plastic window
    ph64 = dict(zip(PHASES, (u * 1e-6 for u in d64.timeFusedPhases(args.kernel_reps))))
    steps64 += args.kernel_reps

    api = None
    if not args.no_api:
        s64a = timed(lambda: d64.run(args.steps, reaction=False), torch) / args.steps
        steps64 += args.steps
        log(f"FP64 per-step API {s64a * 1e3:.3f} ms/step")
        api = {"value": E / s64a, "ms_per_step": s64a * 1e3,
               "step_algorithmic_bytes_per_element": total(api_bytes(n, 8), E, N) / E,
               "step_hbm_frac": total(api_bytes(n, 8), E, N) / s64a / HBM_PEAK}

    mixed = None
    if not args.no_mixed:
        log("building FP32-mixed")
        d32, _ = build_domain(WFX_FP32_MIXED, nx, ny, nz, L, -2.0, stream)
        if args.bs32:
            bs32 = [int(q) for q in args.bs32.split(",")]
            d32.setBlockSizes(bs32)
        else:
            bs32 = d32.autotune(reps=5)
        log(f"FP32-mixed block sizes {'(given)' if args.bs32 else '(autotuned)'} {bs32}")
        d32.runFused(args.pre_steps + args.warmup, reaction=False)
        s32 = timed(lambda: d32.runFused(args.steps, reaction=False), torch) / args.steps
        log(f"FP32-mixed fused {s32 * 1e3:.3f} ms/step")
        ph32 = dict(zip(PHASES, (u * 1e-6 for u in d32.timeFusedPhases(args.kernel_reps))))
        steps32 = args.pre_steps + args.warmup + args.steps + args.kernel_reps
        # accuracy: both paths to err_steps, then one more step with the reaction
        target = max(args.err_steps, steps64, steps32)
        d64.runFused(target - steps64, reaction=False)
        d32.runFused(target - steps32, reaction=False)
        r64 = d64.runFused(1)[0]
        r32 = d32.runFused(1)[0]

        def rel(a, b):
            return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

        mixed = {
            "value": E / s32,
            "ms_per_step": s32 * 1e3,
            "speedup_vs_fp64": s64 / s32,
            "block_sizes": bs32,
            "dtype": "f32 element arithmetic, f64 nodal kinematics",
            "rel_err_vs_fp64": {"displacement": rel(d32.field("u"), d64.field("u")),
                                "reaction_force": abs(r32 - r64) / abs(r64),
                                "eq_plastic_strain": rel(d32.field("pl_strain"), d64.field("pl_strain")),
                                "max_eq_plastic_strain": float(d64.field("pl_strain").max()),
                                "after_steps": target + 1,
                                "reference": "FP64 GPU path (parity-tested against the oracle)"},
            "step_algorithmic_bytes_per_element": total(domain_fused_bytes(d32, 4), E, N) / E,
            "step_hbm_frac": total(domain_fused_bytes(d32, 4), E, N) / s32 / HBM_PEAK,
        }
        mixed["per_kernel"] = per_kernel_table(ph32, domain_fused_bytes(d32, 4), E, N)
        mixed["step_minus_passes_us"] = (s32 - sum(ph32.values())) * 1e6
        log(f"mixed error: {mixed['rel_err_vs_fp64']}")

    kt64 = ph64
    fb64 = domain_fused_bytes(d64, 8)
    dom = max(kt64, key=kt64.get)
    be, bn = fb64[dom]
    dom_bytes = be * E + bn * N
    achieved = dom_bytes / kt64[dom] / 1e9
    step_bytes = total(fb64, E, N)
    if api is not None:
        api["per_kernel"] = per_kernel_table(event_times(api_calls(d64), args.kernel_reps, torch),
                                             api_bytes(n, 8), E, N)

    extensions = None
    if not args.no_extensions:
        blocks = [("fp64", WFX_FP64, (nx, ny, nz, L), bs64)]
        if not args.no_mixed:
            blocks.append(("fp32_mixed", WFX_FP32_MIXED, (nx, ny, nz, L), bs32))
        del d64
        if not args.no_mixed:
            del d32
        extensions = extensions_leg(torch, blocks, 100, log)

    io = None
    if not args.no_io:
        io = io_leg(torch, log)

    large = None
    if not args.no_large and not args.quick:
        large = large_leg(torch, log)

    batch = None
    if not args.no_batch:
        batch = batch_leg(torch, args.batch_steps, log)

    others = None
    if not args.no_configs:
        others = config_legs(torch, log)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(torch, log, with_oracle=not (args.no_cpu or args.quick))

    cpu = None
    if not args.no_cpu and not args.quick:
        log("CPU baseline (oracle)")
        cpu = cpu_baseline()

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(f"{nx}x{ny}x{nz}", {}).get("fp64", {}).get(dom)

    out = {
        "metric": "element-steps/s at 1 GPU (FP64 vs FP32-mixed, stated err); % HBM peak",
        "value": E / s64,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"hexa compression block {nx}x{ny}x{nz} ({E} elements, {N} nodes), "
                               "fused four-kernel FP64 step",
                   "elements": E, "nodes": N, "dt": dt, "parallelism": "single GPU",
                   "block_sizes": bs64,
                   "timed_steps": [args.pre_steps + args.warmup + 1, args.pre_steps + args.warmup + args.steps],
                   "yielded_element_fraction_at_timing": yielded64},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved * 1e9 / HBM_PEAK, "traffic": traffic,
                     "algorithmic_bytes": dom_bytes, "avg_launch_us": kt64[dom] * 1e6,
                     "measured_copy_GBs": copy_bw / 1e9, "frac_of_measured_copy": achieved * 1e9 / copy_bw,
                     "copy_GBs_torch": copy_torch / 1e9, "copy_GBs_stream": copy_stream / 1e9},
        "step_algorithmic_bytes_per_element": step_bytes / E,
        "step_hbm_frac": step_bytes / s64 / HBM_PEAK,
        "per_kernel": per_kernel_table(kt64, fb64, E, N),
        "per_kernel_timing": "HIP events between the four passes of fused runs (wfx_time_fused_phases), "
                             "in the plastic window right after the timed steps",
        # the step time minus the sum of the four event-timed passes (the
        # events between passes add a little, so this can be slightly negative)
        "step_minus_passes_us": (s64 - sum(kt64.values())) * 1e6,
        "fp32_mixed": mixed,
        "api_path_fp64": api,
        "other_configs": others,
        "extensions": extensions,
        "io": io,
        "large_block": large,
        "batched_specimens": batch,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
