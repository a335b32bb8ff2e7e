"""Build recipe of libgalic_hip.so (gfx950) -- invoked by __graft_entry__.build().

Each HIP translation unit is compiled with hipcc for gfx950 only, then linked
into one shared library next to this file.  The FP64 parity path is compiled
with -ffp-contract=off so a*b+c rounds twice, as in the reference's x86-64
build; the FP32 kernels allow contraction.  Packed FP32 instructions are
disabled for all device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgalic_hip.so")

COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-fno-gpu-rdc",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas"]

# translation unit -> extra flags
UNITS = {
    "galic_core.hip": ["-ffp-contract=off"],
    "galic_api.hip": ["-ffp-contract=off"],
    "galic_dropin.hip": ["-ffp-contract=off"],
    "galic_fast.hip": ["-ffp-contract=fast"],
    "galic_tree.hip": ["-ffp-contract=off"],
    "galic_targets.hip": ["-ffp-contract=off"],
    "galic_io.hip": ["-ffp-contract=off"],
    "galic_noalloc.hip": ["-ffp-contract=off"],
    "galic_commit.hip": ["-ffp-contract=off"],
}


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgalic_hip.so")
    return path


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs, procs = [], []
    for unit, extra in UNITS.items():
        src = os.path.join(CSRC, unit)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
        cmd = [cc, *COMMON, *extra, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        # the units compile concurrently (each is one hipcc process)
        procs.append((unit, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for unit, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed.append(unit)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libgalic_hip.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
