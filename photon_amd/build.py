"""Build the gfx950 library photon_amd/libphoton_gas.so in-tree with hipcc.

python -m photon_amd.build     (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libphoton_gas.so")
SOURCES = ["phg.hip", "hierarchy.cpp", "hierarchy_dev.hip"]
HEADERS = ["device.hpp", "kernels_pcg.hpp", "kernels_step.hpp", "kernels_transport.hpp", "kernels_flow.hpp",
           "kernels_leak.hpp", "hierarchy.hpp"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the reference runs with FTZ/DAZ set (src/core/jobs.h:15-17)
    "-fgpu-flush-denormals-to-zero",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


# the C++ host class over the C ABI (no HIP headers: plain g++)
HOST_SRC = os.path.join(HERE, "host", "gas_phase.cpp")
HOST_OUT = os.path.join(HERE, "libphoton_gasphase.so")
# a C++ caller of the host class, run by the GPU tests
CHECK_SRC = os.path.join(HERE, "..", "tests", "cpp", "gas_step_check.cpp")
CHECK_OUT = os.path.join(HERE, "..", "tests", "cpp", "gas_step_check")
CXX = ["g++", "-std=c++20", "-O2", "-Wall", "-Wno-unused-result"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    cmd = [hipcc(), *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-o", OUT + ".tmp",
           *[os.path.join(CSRC, s) for s in SOURCES]]
    _run(cmd, verbose)
    os.replace(OUT + ".tmp", OUT)
    _run([*CXX, "-fPIC", "-shared", "-o", HOST_OUT + ".tmp", HOST_SRC, "-L" + HERE, "-lphoton_gas",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(HOST_OUT + ".tmp", HOST_OUT)
    _run([*CXX, "-o", CHECK_OUT + ".tmp", CHECK_SRC, "-L" + HERE, "-lphoton_gasphase", "-lphoton_gas",
          "-Wl,-rpath,$ORIGIN/../../photon_amd"], verbose)
    os.replace(CHECK_OUT + ".tmp", CHECK_OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
