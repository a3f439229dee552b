"""Builds libpmsim_amd.so (HIP for gfx950) in-tree with hipcc.

    python -m pmsim_amd.build          # or pmsim_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmsim_amd.so")
SOURCES = ["pm_shapes.cpp", "pm_host.cpp", "pm_device.cpp", "pm_solver.cpp", "pm_capi.cpp", "pm_serve.cpp", "pm_kernels.hip", "pm_tangent_mfma.hip", "pm_tangent_hex20.hip", "pm_tangent_hex20_mfma.hip", "pm_linalg.hip", "pm_pcg_batch.hip", "pm_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False, jobs=8):
    os.makedirs(OUT, exist_ok=True)
    objs, procs = [], []
    for s in SOURCES:
        src = os.path.join(SRC, s)
        obj = os.path.join(OUT, s.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-c", src, "-o", obj]
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _wait(procs, verbose)
            procs = []
    _wait(procs, verbose)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def _wait(procs, verbose):
    err = []
    for name, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            err.append(f"--- {name} ---\n{out}")
        elif verbose and out.strip():
            print(f"--- {name} ---\n{out}")
    if err:
        raise RuntimeError("hipcc failed:\n" + "\n".join(err))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
