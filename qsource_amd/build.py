"""Build steps for the engine library (HIP, gfx950) and the CPU checkers.

``build_engine()`` compiles ``qsource_amd/libqgpu.so`` in-tree with hipcc;
``build_oracle()`` compiles ``oracle/liboracle.so`` (test infrastructure);
``build_reference()`` compiles the reference's serial qdyn into
``oracle/_ref`` when the reference sources are present.
"""
from __future__ import annotations

import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qsource_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libqgpu.so")

HIP_SOURCES = ["engine.hip", "kernels_water.hip", "kernels_solute.hip", "kernels_far.hip",
               "kernels_mlp.hip", "mlp_host.hip"]
HOST_SOURCES = ["host/readers.cpp", "host/prep.cpp", "host/mdfiles.cpp"]

# -packed-fp32-ops off: hand-off of FP32 pairs to v_pk_* is disabled;
# -ffp-contract=off: FP64 expressions keep the reference's rounding steps.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout[-6000:]))
    return r.stdout


def build_engine(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in HIP_SOURCES + HOST_SOURCES:
        obj = os.path.join(objdir, os.path.basename(src).replace(".", "_") + ".o")
        flags = list(HIP_FLAGS)
        if src.endswith(".cpp"):
            flags = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]
            cmds.append(["g++"] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        else:
            cmds.append([hipcc] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units compile independently: a few at a time
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, (os.cpu_count() or 2) // 2))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_reference() -> bool:
    if not os.path.isdir("/root/reference/src"):
        return False
    out = os.path.join(ROOT, "oracle", "_ref")
    _run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile.ref"), "OUT=" + out], cwd=ROOT)
    return True


if __name__ == "__main__":
    print(build_engine(verbose=True))
    print(build_oracle())
    print(build_reference())
