"""Build libupf.so (gfx950) in-tree with hipcc.

    python -m unnatural_planets_amd.build          # or build() from Python

The shared library lands next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libupf.so")

SOURCES = ["upf_common.cpp", "upf_field.hip", "upf_mc.hip", "upf_batch.hip", "upf_bake.hip", "upf_nb.hip", "upf_meshpost.hip", "upf_material.hip"]
HEADERS = ["upf_device.hpp", "upf_device_grad.hpp", "upf_device_row.hpp", "upf_internal.hpp", "upf_mc.hpp", "upf_mc_device.hpp", "mc_table.inc",
           os.path.join("..", "..", "include", "upf.h")]

# -ffp-contract=off: every fused multiply-add in the kernels is an explicit fmaf(), so the
# lattice coordinates round exactly as in the CPU restatement.
# -packed-fp32-ops off: the compiler must not form packed FP32 (v_pk_*_f32) instructions.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, extra_flags=()) -> str:
    """Compile libupf.so for gfx950 (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libupf.so")
    errs = [ln for ln in res.stderr.splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln and "target-feature" not in ln]
    if verbose and errs:
        sys.stderr.write("\n".join(errs) + "\n")
    os.replace(tmp, LIB)
    return LIB


ROOT = os.path.dirname(HERE)
CALLER_SRC = os.path.join(ROOT, "tests", "integration", "mesh_generate_generic.cpp")
CALLER_BIN = os.path.join(ROOT, "tests", "integration", "mesh_generate_generic")


def build_caller(verbose: bool = False) -> str:
    """Compile the integration guide's C++ caller (host code against include/upf.h,
    linked with libupf.so) next to its source."""
    cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", CALLER_BIN + ".tmp", CALLER_SRC,
           "-L", HERE, "-lupf", "-Wl,-rpath,$ORIGIN/../../unnatural_planets_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building the integration caller")
    os.replace(CALLER_BIN + ".tmp", CALLER_BIN)
    return CALLER_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_caller(verbose=True))
