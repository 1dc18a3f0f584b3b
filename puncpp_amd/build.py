"""Build recipe for libpunc_gpu.so (HIP kernels for gfx950 + host setup + C ABI).

    python -m puncpp_amd.build        # or puncpp_amd.build.build()

hipcc compiles the device code for gfx950 only.  -ffp-contract=off keeps every
floating-point expression exactly as written (no FMA contraction), which is
what makes positions, velocities and cell decisions bit-identical to the
serial CPU path.  Packed-FP32 instruction selection is disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpunc_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
SOURCES = ["kernels.hip", "capi.hip", "adjoint.hip", "field.hip", "mesh_host.cpp", "amg.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC] + DEVICE + COMMON + ["-c", path, "-o", obj]
        else:
            cmd = [HIPCC, "-x", "c++"] + COMMON + ["-c", path, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    # the sources compile independently: one compiler process each
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
