"""Builds libnbx.so (the HIP simulation core) in-tree for gfx950.

Flags that the numerics depend on:
  -ffp-contract=off                 no FMA contraction: every float op rounds as written
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division (the model RHS divides by tau)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libnbx.so")
ARCH = os.environ.get("NBX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["nbx_core.hip", "nbx_step_kernels.hip", "nbx_persistent.hip", "nbx_build.hip", "nbx_batch.hip", "nbx_adjoint.hip",
           "nbx_nbnet.hip", "nbx_monitor.hip", "nbx_replay_stdp.hip", "nbx_currents.hip",
           "nbx_small.hip"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libnbx.so (always recompiles)."""
    objs = []
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    failed = []
    for p, src in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(f"--- {src}\n{text}")
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


NATIVE_TESTS = ["dense_runtime_smoke", "dense_runtime_parity"]


def build_native_tests() -> list:
    """g++ builds of the C++ checks under tests/native (host code only): each links the reference-
    interface runtime classes (shim/npgr_runtime.cpp) over libnbx.so, the HIP runtime for its own
    device buffers, and the CPU oracle as the checker."""
    out = []
    src_dir = os.path.join(ROOT, "tests", "native")
    oracle = os.path.join(ROOT, "oracle")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    for name in NATIVE_TESTS:
        exe = os.path.join(src_dir, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
               "-I", os.path.join(PKG, "shim", "npgr"), "-I", oracle, "-I", os.path.join(rocm, "include"),
               os.path.join(src_dir, name + ".cpp"), os.path.join(PKG, "shim", "npgr_runtime.cpp"), "-o", exe,
               "-L", PKG, "-lnbx", "-L", oracle, "-lorc", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
               "-Wl,-rpath," + PKG, "-Wl,-rpath," + oracle, "-Wl,-rpath,$ORIGIN/../../neuronbridge_amd",
               "-Wl,-rpath,$ORIGIN/../../oracle", "-Wl,-rpath," + os.path.join(rocm, "lib")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {name}:\n{r.stdout}{r.stderr}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
