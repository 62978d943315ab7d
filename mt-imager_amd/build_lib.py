"""Builds libmtimager_amd.so (HIP kernels + C-ABI + C++ host layer) for gfx950.

Run directly (`python mt-imager_amd/build_lib.py`) or through
__graft_entry__.build().  Always recompiles.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmtimager_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable",
          # packed FP32 instructions are disabled for all device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]
# Bit-exact FP64 binning and kernel generation: no FMA contraction.
EXACT = {"mti_bin.hip", "mti_wkernel.hip", "mti_degrid.hip"}


def _sources():
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    host = sorted(f for f in os.listdir(HOST) if f.endswith(".cpp")) if os.path.isdir(HOST) else []
    return [os.path.join(CSRC, s) for s in srcs] + [os.path.join(HOST, h) for h in host]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    flags = list(COMMON)
    if os.path.basename(src) in EXACT:
        flags.append("-ffp-contract=off")
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + flags + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    return src, obj, p.returncode, out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    failed = False
    for src, obj, rc, out in results:
        if rc != 0:
            failed = True
            sys.stderr.write(f"[build] FAILED {os.path.basename(src)}\n{out}\n")
        elif verbose and out.strip():
            sys.stderr.write(f"[build] {os.path.basename(src)}:\n{out}\n")
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
          [r[1] for r in results]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    build_probe()
    return LIB


PROBE_SRC = os.path.join(HERE, "probe", "mti_probe.hip")
PROBE_LIB = os.path.join(HERE, "libmti_probe.so")


def build_probe():
    """The benchmark's achievable-peak probes (stream copy, FP64 FMA loop)."""
    tmp = PROBE_LIB + ".tmp"
    flags = list(COMMON)
    i = flags.index("-I")
    del flags[i:i + 2]
    cmd = [HIPCC] + flags + ["-shared", "-o", tmp, PROBE_SRC]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("probe build failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


if __name__ == "__main__":
    print(build(verbose=True))
