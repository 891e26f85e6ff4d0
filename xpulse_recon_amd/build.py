"""Build the HIP library of the hot path in-tree (gfx950).

``python -m xpulse_recon_amd.build`` compiles ``csrc/xpr_kernels.hip`` into
``xpulse_recon_amd/lib/libxpr_hip.so`` with hipcc.  Every build compiles from
source, whether or not a library is already there.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libxpr_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: only explicit fmaf calls fuse (the canonical arithmetic).
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 in generated code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


BIN_DIR = os.path.join(HERE, "bin")
CLI = os.path.join(BIN_DIR, "XPRecon")
HOST_SRCS = ["json.cpp", "params.cpp", "io.cpp", "recon.cpp", "capi.cpp", "serve.cpp"]


def build_hip_lib() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", f) for f in ("xpr_kernels.hip", "geometry.cpp", "synth.hip")]
    srcs += [os.path.join(HERE, "csrc", "host", f) for f in HOST_SRCS]
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs, "-ldl", "-lpthread"])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_cli() -> str:
    """The XPRecon command line, linked to the library with an $ORIGIN rpath."""
    os.makedirs(BIN_DIR, exist_ok=True)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-o", tmp,
          os.path.join(HERE, "csrc", "host", "main.cpp"), "-L" + LIB_DIR, "-lxpr_hip",
          "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"])
    os.replace(tmp, CLI)
    return CLI


def build_all() -> None:
    build_hip_lib()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
