"""Build the HIP library in-tree for gfx950 (hipcc cross-compiles without a GPU)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "lmfhd.hip")
OUT = os.path.join(HERE, "liblmfhd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: hipcc must not emit v_pk_{add,mul,fma}_f32 for this target image.
# -ffp-contract=off: every a*b+c rounds twice, as written.  With contraction the compiler may fuse
# a functor's products and sums differently in different kernels that inline it (the launch
# forms of one functor, the lock-step batched form), and results that must be the same bits in
# every launch form would not be; the kernels are memory-bound, so fused multiply-adds buy
# nothing here.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_library(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", OUT + ".tmp", SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for " + SRC)
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


PROBE_SRC = os.path.join(HERE, "csrc", "stream_probe.hip")
PROBE = os.path.join(HERE, "stream_probe")


def build_probe() -> str:
    """The streaming-bandwidth probe bench.py reports beside the roofline (a gfx950 executable)."""
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-o", PROBE + ".tmp", PROBE_SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for " + PROBE_SRC)
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


HARNESS_SRC = os.path.join(os.path.dirname(HERE), "tests", "abi", "fab_harness.c")
HARNESS = os.path.join(os.path.dirname(HERE), "tests", "abi", "fab_harness")


def build_harness() -> str:
    """The C caller of the ABI used by tests/test_gpu_c_harness.py (host code, gcc), linked to
    the in-tree library through an rpath relative to itself."""
    cmd = ["gcc", "-O1", "-Wall", "-o", HARNESS, HARNESS_SRC, "-L" + HERE, "-llmfhd",
           "-Wl,-rpath,$ORIGIN/../../lowmachfhd_amd", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed for " + HARNESS_SRC)
    return HARNESS


if __name__ == "__main__":
    print(build_library(verbose=True))
