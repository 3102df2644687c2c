"""Builds the in-tree HIP step library multiphase-sph-solver_amd/_build/libsphg.so for gfx950.

Run directly (python build.py) or through __graft_entry__.build(). FP contraction is off for all
device and host code: the FP64 kernels are meant to reproduce the reference's CPU arithmetic.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libsphg.so")
SOLVE = os.path.join(OUT_DIR, "sphg_solve")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def kernel_resources(remarks: str) -> dict:
    """Per-kernel VGPRs / scratch bytes per lane / occupancy from the compiler's resource remarks."""
    out, cur = {}, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and cur:
            out[cur][m.group(1).split()[0]] = int(m.group(2))
    return out


def check_no_scratch(remarks: str) -> None:
    """Every device function must stay in registers: a kernel that spills to per-lane scratch
    memory (a stack array whose address escapes, an oversized unrolled body) runs several times
    slower, so such a build is refused."""
    res = kernel_resources(remarks)
    if not res:
        raise RuntimeError("build: no kernel resource remarks found")
    bad = {k: v["ScratchSize"] for k, v in res.items() if v.get("ScratchSize", 0) > 0}
    if bad:
        raise RuntimeError("build: kernels use scratch memory: " + ", ".join(f"{k} {b} B/lane" for k, b in bad.items()))


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_cpp = os.path.join(OUT_DIR, "sph_init.o")
    obj_kl = os.path.join(OUT_DIR, "sph_kl_host.o")
    obj_io = os.path.join(OUT_DIR, "sph_io.o")
    obj_drv = os.path.join(OUT_DIR, "sph_driver.o")
    obj_par = os.path.join(OUT_DIR, "sph_particles.o")
    obj_hip = os.path.join(OUT_DIR, "sphg_capi.o")
    host = ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-c"]
    cmds = [
        host + [os.path.join(CSRC, "sph_init.cpp"), "-o", obj_cpp],
        host + [os.path.join(CSRC, "sph_kl_host.cpp"), "-o", obj_kl],
        host + [os.path.join(CSRC, "sph_io.cpp"), "-o", obj_io],
        host + [os.path.join(CSRC, "sph_driver.cpp"), "-o", obj_drv],
        host + [os.path.join(CSRC, "sph_particles.cpp"), "-o", obj_par],
        [HIPCC] + HIP_FLAGS + ["-Rpass-analysis=kernel-resource-usage", "-c", os.path.join(CSRC, "sphg_capi.hip"),
                               "-o", obj_hip],
        [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", obj_hip, obj_cpp, obj_kl, obj_io, obj_drv,
                                                                                 obj_par],
    ]
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"build failed: {' '.join(c[:3])} ...")
        if any("kernel-resource-usage" in a for a in c):
            check_no_scratch(r.stderr)
            # per-kernel VGPRs / occupancy of this build, for reading beside the profiles
            with open(os.path.join(OUT_DIR, "kernel_resources.json"), "w") as f:
                json.dump(kernel_resources(r.stderr), f, indent=1, sort_keys=True)
    os.replace(LIB + ".tmp", LIB)
    # the project runner executable, linked against the library beside it
    r = subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "sphg_main.cpp"), "-o", SOLVE,
                        "-L" + OUT_DIR, "-lsphg", "-Wl,-rpath,$ORIGIN"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: sphg_solve")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
