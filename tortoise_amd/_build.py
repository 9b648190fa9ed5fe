"""Build recipe of libtortoise_gpu.so (hipcc, gfx950, in-tree)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtortoise_gpu.so")
SOURCES = ["tg_context.hip", "tg_assemble.hip", "tg_plain.hip", "tg_stack.hip", "tg_mixed.hip",
           "tg_dp5.hip", "tg_adjoint.hip", "tg_ee.hip", "tg_tuple.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: hipcc must not form v_pk_{add,mul,fma}_f32 from float code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtortoise_gpu.so")
    os.replace(tmp, LIB)
    return LIB


def build_host_example(verbose: bool = False) -> str:
    """The C++ host-layer example (tortoise_amd/host), linked against the
    library just built: its GPU tests run the binary."""
    cmd = ["make", "-B", "-C", os.path.join(HERE, "host")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the host-layer example failed")
    return os.path.join(HERE, "host", "examples", "thermalisation_square")


if __name__ == "__main__":
    print(build(verbose=True))
