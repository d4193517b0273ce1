"""Compile libmpm_hip.so (HIP kernels + C ABI) for gfx950 with hipcc.

Every translation unit is compiled separately (in parallel) into
``mpm_cuda_amd/_build/`` and linked into ``mpm_cuda_amd/libmpm_hip.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpm_hip.so")
ARCH = os.environ.get("MPM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mpm_api.hip", "mpm_step.hip", "mpm_fast.hip", "mpm_batch.hip", "mpm2d.hip", "mpm_mlp.hip", "mpm_probe.hip",
           "scene.cpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libmpm_hip.so)")


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        cmd = [cc, *HIP_FLAGS, "-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [cc, "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s} ---\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, LIB)
    # command-line driver over the C++ mirror header, linked against the library
    cli = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "mpm_run.cpp"),
           "-o", os.path.join(HERE, "mpm_run"), f"-L{HERE}", "-lmpm_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cli, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mpm_run build failed:\n{r.stdout}")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
