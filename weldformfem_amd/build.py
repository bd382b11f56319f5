"""Build recipe of the wfx HIP library (weldformfem_amd/libwfx.so) for gfx950.

Plain hipcc/g++ invocations, no build system.  The FP64 per-step kernel unit is
built with -ffp-contract=off so its expressions round like the reference's
CPU code; the fused FP64 unit and the FP32-mixed unit allow contraction.  Every HIP unit is built
without packed-FP32 instructions.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libwfx.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("WFX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
UNITS = [
    ("wfx_k64.hip", ["-ffp-contract=off"]),
    ("wfx_f64.hip", ["-ffp-contract=fast"]),
    ("wfx_k32.hip", []),
    ("wfx_api.hip", ["-ffp-contract=off"]),
    ("wfx_adjoint.hip", ["-ffp-contract=off"]),
    ("wfx_frames.hip", []),
]


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build(verbose=False, jobs=None):
    hipcc = _hipcc()
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmds.append([hipcc] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("wfx_mesh.cpp", "wfx_io.cpp", "wfx_cluster.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the units are independent: compile them concurrently
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 8)
    pending, running = list(cmds), []
    while pending or running:
        while pending and len(running) < jobs:
            cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd))
            running.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                  text=True)))
        cmd, proc = running.pop(0)
        out, _ = proc.communicate()
        if proc.returncode != 0:
            for _, p in running:
                p.wait()
            sys.stderr.write(out)
            raise RuntimeError("build step failed: " + " ".join(cmd))
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    # C++ driver over the header-only Domain_d mirror (host code only)
    inc = os.path.join(os.path.dirname(HERE), "include")
    _run(["g++", "-O2", "-std=c++17", "-I" + inc, os.path.join(HERE, "cpp", "cube_compression.cpp"),
          "-L" + HERE, "-lwfx", "-Wl,-rpath,$ORIGIN", "-o", os.path.join(HERE, "wfx_cube")])
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
