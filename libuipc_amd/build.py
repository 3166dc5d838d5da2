"""Builds the package's native libraries in-tree.

  libuipc_amd/libmi_linsys.so  HIP kernels + C ABI (include/mi_linsys.h), gfx950 only
  libuipc_amd/libmi_synth.so   host C++ synthetic-system generator (include/mi_synth.h)

Objects go to libuipc_amd/_build/.  `python -m libuipc_amd.build` or build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
ROOT = os.path.dirname(PKG)

HIP_SOURCES = ["convert.hip", "spmv.hip", "vec.hip", "batch.hip", "precond.hip", "mas.hip", "mas_init.hip", "partition.hip", "subsystem.hip", "producers.hip", "producers_abd.hip", "producers_shell.hip", "warmstart.hip", "linsys.hip"]
HOST_SOURCES = ["mtx.cpp"]  # linked into libmi_linsys.so, no device code
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # hardware FP64/FP32 global atomics (global_atomic_add_f64)
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> dict[str, str]:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    for src in HOST_SOURCES:
        o = os.path.join(OBJ, src.replace(".cpp", ".o"))
        objs.append(o)
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
                     "-c", os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 4)) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(PKG, "libmi_linsys.so")
    tmp = lib + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, lib)

    synth = os.path.join(PKG, "libmi_synth.so")
    tmp = synth + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          f"-I{os.path.join(ROOT, 'include')}", "-o", tmp, os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, synth)
    if verbose:
        print(f"built {lib}\nbuilt {synth}")
    return {"linsys": lib, "synth": synth}


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
